// scene.cpp — host scene model, binned-SAH BVH builder, camera and the
// procedural scenes the renderer is exercised on.
//
// Follows rustracer-core: scene.rs (Mesh/Scene/flatten/bounds/cornell_box/
// dark_room/create_quad/create_cube), bvh.rs (AABB, BVHBuilder), material.rs
// (constructors, to_gpu), camera.rs (to_uniform, frame_scene, orbit, zoom, pan)
// and rustracer-renderer/src/lib.rs:1231-1282 (light_to_gpu). Arithmetic is
// float32 in glam's evaluation order, built without FMA contraction.
#include "scene.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <stdexcept>

namespace rt {

// ───────────────────────────── BVH (bvh.rs:120-336) ─────────────────────────────

namespace {

constexpr int kSahBins = 12;      // bvh.rs:231
constexpr size_t kMaxLeaf = 4;    // bvh.rs:131

struct Bin {
    AABB bounds;
    size_t count = 0;
};

// Rust `f32 as usize`: truncation, saturating, NaN -> 0.
inline size_t f32_as_usize(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 1.8446744e19f) return SIZE_MAX;
    return (size_t)v;
}

class SahBuilder {
  public:
    SahBuilder(const std::vector<Vec3> &c, const std::vector<AABB> &b) : cent_(c), box_(b) {}

    BVH build() const {
        BVH out;
        out.prim_indices.resize(cent_.size());
        for (size_t i = 0; i < cent_.size(); i++) out.prim_indices[i] = (uint32_t)i;
        subtree(out.nodes, out.prim_indices, 0, out.prim_indices.size());
        reverse(out.nodes);
        return out;
    }

  private:
    const std::vector<Vec3> &cent_;
    const std::vector<AABB> &box_;

    // bvh.rs:153-215 — children are emitted before their parent.
    uint32_t subtree(std::vector<rt_bvh_node> &nodes, std::vector<uint32_t> &prims, size_t start,
                     size_t end) const {
        const size_t count = end - start;
        if (count == 0) return 0;
        AABB bb;
        for (size_t i = start; i < end; i++) bb = bb.unite(box_[prims[i]]);
        int axis = 0;
        size_t left_count = 0;
        const bool split = count > kMaxLeaf && best_split(prims, start, end, bb, axis, left_count);
        rt_bvh_node node;
        std::memset(&node, 0, sizeof(node));
        node.bbox_min[0] = bb.min.x; node.bbox_min[1] = bb.min.y; node.bbox_min[2] = bb.min.z;
        node.bbox_max[0] = bb.max.x; node.bbox_max[1] = bb.max.y; node.bbox_max[2] = bb.max.z;
        if (!split) {
            node.left_first = (uint32_t)start;
            node.right_child = 0;
            node.prim_count = (uint32_t)count;
            nodes.push_back(node);
            return (uint32_t)(nodes.size() - 1);
        }
        const size_t mid = start + left_count;
        // select_nth_unstable_by on the centroid axis: the left_count smallest go left.
        std::nth_element(prims.begin() + (ptrdiff_t)start, prims.begin() + (ptrdiff_t)mid,
                         prims.begin() + (ptrdiff_t)end,
                         [&](uint32_t a, uint32_t b) { return cent_[a][axis] < cent_[b][axis]; });
        const uint32_t left_idx = subtree(nodes, prims, start, mid);
        const uint32_t right_idx = subtree(nodes, prims, mid, end);
        node.left_first = left_idx;
        node.right_child = right_idx;
        node.prim_count = 0;
        nodes.push_back(node);
        return (uint32_t)(nodes.size() - 1);
    }

    // bvh.rs:224-319 — binned SAH over centroid positions, real boxes per bin.
    bool best_split(const std::vector<uint32_t> &prims, size_t start, size_t end, const AABB &bb, int &best_axis,
                    size_t &best_left) const {
        const float bb_sa = bb.surface_area();
        if (bb_sa <= 0.0f) return false;
        const float inv_bb_sa = 1.0f / bb_sa;
        Vec3 cmin{3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
        Vec3 cmax{-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
        for (size_t i = start; i < end; i++) {
            cmin = vmin(cmin, cent_[prims[i]]);
            cmax = vmax(cmax, cent_[prims[i]]);
        }
        best_axis = 0;
        best_left = 0;
        float best_cost = INFINITY;
        for (int axis = 0; axis < 3; axis++) {
            const float range = cmax[axis] - cmin[axis];
            if (range <= 0.0f) continue;
            const float scale = (float)kSahBins / range;
            Bin bins[kSahBins];
            for (size_t i = start; i < end; i++) {
                const uint32_t p = prims[i];
                size_t bin = f32_as_usize((cent_[p][axis] - cmin[axis]) * scale);
                if (bin >= (size_t)kSahBins) bin = kSahBins - 1;
                bins[bin].count++;
                bins[bin].bounds = bins[bin].bounds.unite(box_[p]);
            }
            float left_area[kSahBins], right_area[kSahBins];
            size_t left_n[kSahBins], right_n[kSahBins];
            AABB acc;
            size_t n = 0;
            for (int b = 0; b < kSahBins; b++) {
                acc = acc.unite(bins[b].bounds);
                n += bins[b].count;
                left_area[b] = acc.surface_area();
                left_n[b] = n;
            }
            acc = AABB();
            n = 0;
            for (int b = kSahBins - 1; b >= 0; b--) {
                acc = acc.unite(bins[b].bounds);
                n += bins[b].count;
                right_area[b] = acc.surface_area();
                right_n[b] = n;
            }
            for (int b = 0; b < kSahBins - 1; b++) {
                const size_t lc = left_n[b], rc = right_n[b + 1];
                if (lc == 0 || rc == 0) continue;
                const float cost = 1.0f + (left_area[b] * (float)lc + right_area[b + 1] * (float)rc) * inv_bb_sa;
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_left = lc;
                }
            }
        }
        const float leaf_cost = (float)(end - start);
        return best_cost < leaf_cost;
    }

    // bvh.rs:322-335
    static void reverse(std::vector<rt_bvh_node> &nodes) {
        const size_t n = nodes.size();
        if (n <= 1) return;
        std::reverse(nodes.begin(), nodes.end());
        for (auto &node : nodes) {
            if (node.prim_count == 0) {
                node.left_first = (uint32_t)(n - 1 - node.left_first);
                node.right_child = (uint32_t)(n - 1 - node.right_child);
            }
        }
    }
};

}  // namespace

BVH build_bvh(const std::vector<Vec3> &centroids, const std::vector<AABB> &boxes) {
    return SahBuilder(centroids, boxes).build();
}

uint32_t bvh_depth(const BVH &bvh) {
    if (bvh.nodes.empty()) return 0;
    uint32_t best = 0;
    std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 1u}};
    while (!stack.empty()) {
        auto [idx, d] = stack.back();
        stack.pop_back();
        best = std::max(best, d);
        const rt_bvh_node &n = bvh.nodes[idx];
        if (n.prim_count == 0) {
            stack.push_back({n.left_first, d + 1});
            stack.push_back({n.right_child, d + 1});
        }
    }
    return best;
}

// ───────────────────────────── Scene (scene.rs) ─────────────────────────────

size_t Scene::triangle_count() const {
    size_t n = 0;
    for (const auto &m : meshes) n += m.triangle_count();
    return n;
}

bool Scene::bounds(Vec3 &bmin, Vec3 &bmax) const {
    AABB bb;
    for (const auto &m : meshes)
        for (const auto &p : m.positions) bb.extend(p);
    if (bb.min.x > bb.max.x) return false;
    bmin = bb.min;
    bmax = bb.max;
    return true;
}

std::vector<rt_triangle> Scene::flatten_triangles() const {
    std::vector<rt_triangle> tris;
    tris.reserve(triangle_count());
    for (const auto &m : meshes) {
        for (size_t t = 0; t < m.triangle_count(); t++) {
            rt_triangle g;
            std::memset(&g, 0, sizeof(g));
            const uint32_t id[3] = {m.indices[3 * t], m.indices[3 * t + 1], m.indices[3 * t + 2]};
            float *vs[3] = {g.v0, g.v1, g.v2};
            float *ns[3] = {g.n0, g.n1, g.n2};
            float *us[3] = {g.uv0, g.uv1, g.uv2};
            for (int k = 0; k < 3; k++) {
                const Vec3 p = id[k] < m.positions.size() ? m.positions[id[k]] : Vec3{0.f, 0.f, 0.f};
                const Vec3 n = id[k] < m.normals.size() ? m.normals[id[k]] : Vec3{0.f, 1.f, 0.f};
                const Vec2 uv = id[k] < m.texcoords.size() ? m.texcoords[id[k]] : Vec2{0.f, 0.f};
                vs[k][0] = p.x; vs[k][1] = p.y; vs[k][2] = p.z;
                ns[k][0] = n.x; ns[k][1] = n.y; ns[k][2] = n.z;
                us[k][0] = uv.x; us[k][1] = uv.y;
            }
            g.material_id = m.material_id;
            tris.push_back(g);
        }
    }
    return tris;
}

BVH Scene::build_bvh() const {
    std::vector<Vec3> centroids;
    std::vector<AABB> boxes;
    centroids.reserve(triangle_count());
    boxes.reserve(triangle_count());
    for (const auto &m : meshes) {
        for (size_t t = 0; t < m.triangle_count(); t++) {
            const Vec3 p0 = m.positions[m.indices[3 * t]];
            const Vec3 p1 = m.positions[m.indices[3 * t + 1]];
            const Vec3 p2 = m.positions[m.indices[3 * t + 2]];
            centroids.push_back((p0 + p1 + p2) / 3.0f);
            AABB b;
            b.extend(p0);
            b.extend(p1);
            b.extend(p2);
            boxes.push_back(b);
        }
    }
    return rt::build_bvh(centroids, boxes);
}

rt_light light_to_gpu(const rt_light_desc &l) {
    rt_light g;
    std::memset(&g, 0, sizeof(g));
    g.kind = l.kind;
    g.intensity = l.intensity;
    switch (l.kind) {
        case RT_LIGHT_POINT:
        case RT_LIGHT_DIRECTIONAL:
            for (int k = 0; k < 3; k++) g.data[k] = l.vec_a[k];
            for (int k = 0; k < 3; k++) g.color[k] = l.color[k];
            break;
        case RT_LIGHT_AREA:
            g.data[0] = (float)l.triangle_index;
            for (int k = 0; k < 3; k++) g.color[k] = l.color[k];
            break;
        case RT_LIGHT_ENVIRONMENT:
            for (int k = 0; k < 3; k++) g.data[k] = l.vec_b[k];  // horizon
            for (int k = 0; k < 3; k++) g.color[k] = l.vec_a[k]; // zenith
            break;
        default:
            break;
    }
    return g;
}

std::vector<rt_light> Scene::lights_gpu() const {
    std::vector<rt_light> out;
    for (const auto &l : lights) out.push_back(light_to_gpu(l));
    return out;
}

// ───────────────────────────── materials (material.rs:101-161) ─────────────────────────────

static rt_material material_base(uint32_t kind, Vec3 albedo, float roughness, float metallic, float ior,
                                 Vec3 emissive) {
    rt_material m;
    std::memset(&m, 0, sizeof(m));
    m.albedo[0] = albedo.x; m.albedo[1] = albedo.y; m.albedo[2] = albedo.z; m.albedo[3] = 1.0f;
    m.emissive[0] = emissive.x; m.emissive[1] = emissive.y; m.emissive[2] = emissive.z;
    m.roughness = roughness;
    m.metallic = metallic;
    m.ior = ior;
    m.kind = kind;
    m.flags = 0;
    m.tex_albedo = m.tex_mr = m.tex_emissive = m.tex_normal = RT_TEX_NONE;
    return m;
}
rt_material material_lambertian(Vec3 albedo) {
    return material_base(RT_MAT_LAMBERTIAN, albedo, 1.0f, 0.0f, 1.5f, {0.f, 0.f, 0.f});
}
rt_material material_metal(Vec3 albedo, float roughness) {
    return material_base(RT_MAT_METAL, albedo, std::fmin(std::fmax(roughness, 0.001f), 1.0f), 1.0f, 1.5f,
                         {0.f, 0.f, 0.f});
}
rt_material material_dielectric(float ior, Vec3 albedo) {
    return material_base(RT_MAT_DIELECTRIC, albedo, 0.0f, 0.0f, ior, {0.f, 0.f, 0.f});
}
rt_material material_emissive(Vec3 color, float intensity) {
    return material_base(RT_MAT_EMISSIVE, color, 0.5f, 0.0f, 1.5f, color * intensity);
}

// ───────────────────────────── geometry helpers (scene.rs:463-505) ─────────────────────────────

void add_quad(Scene &s, Vec3 v0, Vec3 v1, Vec3 v2, Vec3 v3, uint32_t material) {
    Mesh m;
    m.positions = {v0, v1, v2, v3};
    const Vec3 n = normalize(cross(v1 - v0, v2 - v0));
    m.normals = {n, n, n, n};
    m.texcoords = {{0.f, 0.f}, {1.f, 0.f}, {1.f, 1.f}, {0.f, 1.f}};
    m.indices = {0, 1, 2, 0, 2, 3};
    m.material_id = material;
    s.meshes.push_back(std::move(m));
}

void add_cube(Scene &s, Vec3 lo, Vec3 hi, uint32_t material) {
    const Vec3 f[6][4] = {
        {{lo.x, lo.y, lo.z}, {lo.x, hi.y, lo.z}, {lo.x, hi.y, hi.z}, {lo.x, lo.y, hi.z}},  // -X
        {{hi.x, lo.y, hi.z}, {hi.x, hi.y, hi.z}, {hi.x, hi.y, lo.z}, {hi.x, lo.y, lo.z}},  // +X
        {{lo.x, lo.y, hi.z}, {hi.x, lo.y, hi.z}, {hi.x, lo.y, lo.z}, {lo.x, lo.y, lo.z}},  // -Y
        {{lo.x, hi.y, lo.z}, {hi.x, hi.y, lo.z}, {hi.x, hi.y, hi.z}, {lo.x, hi.y, hi.z}},  // +Y
        {{hi.x, lo.y, lo.z}, {hi.x, hi.y, lo.z}, {lo.x, hi.y, lo.z}, {lo.x, lo.y, lo.z}},  // -Z
        {{lo.x, lo.y, hi.z}, {lo.x, hi.y, hi.z}, {hi.x, hi.y, hi.z}, {hi.x, lo.y, hi.z}},  // +Z
    };
    for (int i = 0; i < 6; i++) add_quad(s, f[i][0], f[i][1], f[i][2], f[i][3], material);
}

static uint32_t add_material(Scene &s, const rt_material &m) {
    s.materials.push_back(m);
    return (uint32_t)(s.materials.size() - 1);
}

static void add_area_light_quad(Scene &s, Vec3 v0, Vec3 v1, Vec3 v2, Vec3 v3, Vec3 radiance) {
    const uint32_t mat = add_material(s, material_emissive(radiance, 1.0f));
    const uint32_t first = (uint32_t)s.triangle_count();
    add_quad(s, v0, v1, v2, v3, mat);
    for (uint32_t k = 0; k < 2; k++) {
        rt_light_desc l{};
        l.kind = RT_LIGHT_AREA;
        l.triangle_index = first + k;
        l.color[0] = radiance.x; l.color[1] = radiance.y; l.color[2] = radiance.z;
        l.intensity = 1.0f;
        s.lights.push_back(l);
    }
}

static void add_environment(Scene &s, Vec3 zenith, Vec3 horizon, float intensity) {
    rt_light_desc l{};
    l.kind = RT_LIGHT_ENVIRONMENT;
    l.vec_a[0] = zenith.x; l.vec_a[1] = zenith.y; l.vec_a[2] = zenith.z;
    l.vec_b[0] = horizon.x; l.vec_b[1] = horizon.y; l.vec_b[2] = horizon.z;
    l.intensity = intensity;
    s.lights.push_back(l);
}

static void add_point_light(Scene &s, Vec3 pos, Vec3 color, float intensity) {
    rt_light_desc l{};
    l.kind = RT_LIGHT_POINT;
    l.vec_a[0] = pos.x; l.vec_a[1] = pos.y; l.vec_a[2] = pos.z;
    l.color[0] = color.x; l.color[1] = color.y; l.color[2] = color.z;
    l.intensity = intensity;
    s.lights.push_back(l);
}

// Box of half-extent h rotated by `angle` about +Y, standing on y = y0, centred at (cx, cz).
static void add_rotated_box(Scene &s, float cx, float cz, float y0, Vec3 size, float angle, uint32_t mat) {
    const float c = std::cos(angle), sn = std::sin(angle);
    auto P = [&](float x, float y, float z) {
        return Vec3{cx + c * x + sn * z, y0 + y, cz - sn * x + c * z};
    };
    const float hx = size.x * 0.5f, hz = size.z * 0.5f, hy = size.y;
    const Vec3 lo[4] = {P(-hx, 0, -hz), P(hx, 0, -hz), P(hx, 0, hz), P(-hx, 0, hz)};
    const Vec3 up[4] = {P(-hx, hy, -hz), P(hx, hy, -hz), P(hx, hy, hz), P(-hx, hy, hz)};
    add_quad(s, up[0], up[3], up[2], up[1], mat);  // top, facing +Y
    add_quad(s, lo[0], lo[1], lo[2], lo[3], mat);  // bottom, facing -Y
    for (int i = 0; i < 4; i++) {
        const int j = (i + 1) % 4;
        add_quad(s, lo[i], up[i], up[j], lo[j], mat);
    }
}

// ── procedural surfaces ──

static inline uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
static inline float lattice(int32_t ix, int32_t iy, int32_t iz, uint32_t seed) {
    uint32_t h = hash_u32((uint32_t)ix * 73856093u ^ hash_u32((uint32_t)iy * 19349663u ^
                          hash_u32((uint32_t)iz * 83492791u ^ seed)));
    return (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;  // [-1, 1)
}
static inline float fade(float t) { return t * t * (3.0f - 2.0f * t); }
// Trilinear value noise with smoothstep fade, range [-1, 1].
static float value_noise3(float x, float y, float z, uint32_t seed) {
    const float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    const float tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
    float v[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) v[a][b][c] = lattice(ix + a, iy + b, iz + c, seed);
    auto L = [](float a, float b, float t) { return a + (b - a) * t; };
    const float x00 = L(v[0][0][0], v[1][0][0], tx), x10 = L(v[0][1][0], v[1][1][0], tx);
    const float x01 = L(v[0][0][1], v[1][0][1], tx), x11 = L(v[0][1][1], v[1][1][1], tx);
    return L(L(x00, x10, ty), L(x01, x11, ty), tz);
}
static float fbm3(float x, float y, float z, uint32_t seed, int octaves) {
    float sum = 0.f, amp = 0.5f, f = 1.0f;
    for (int o = 0; o < octaves; o++) {
        sum += amp * value_noise3(x * f, y * f, z * f, seed + 1013u * (uint32_t)o);
        amp *= 0.5f;
        f *= 2.0f;
    }
    return sum;
}

// Geodesic sphere: each icosahedron face split into freq^2 triangles
// (20 * freq^2 triangles). disp(dir) gives the radial displacement; normals
// follow the displaced surface. Each icosahedron face becomes its own mesh
// when `face_material` is non-null (one material per face).
template <class Disp>
static void add_geodesic(Scene &s, Vec3 center, float radius, int freq, Disp disp,
                         const std::vector<uint32_t> &face_material) {
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    const Vec3 iv[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int ifc[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                            {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                            {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                            {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    auto surf = [&](Vec3 d) { return center + d * (radius * (1.0f + disp(d))); };
    for (int f = 0; f < 20; f++) {
        Mesh m;
        m.material_id = face_material[(size_t)f % face_material.size()];
        const Vec3 a = iv[ifc[f][0]], b = iv[ifc[f][1]], c = iv[ifc[f][2]];
        // vertex (i, j) with i + j <= freq
        auto vid = [&](int i, int j) { return (uint32_t)(i * (freq + 1) - i * (i - 1) / 2 + j); };
        for (int i = 0; i <= freq; i++) {
            for (int j = 0; j <= freq - i; j++) {
                const float u = (float)i / freq, v = (float)j / freq;
                const Vec3 p = a * (1.0f - u - v) + b * u + c * v;
                const Vec3 d = normalize(p);
                const Vec3 pos = surf(d);
                // normal from the displaced surface around d
                const Vec3 up = std::fabs(d.y) < 0.9f ? Vec3{0, 1, 0} : Vec3{1, 0, 0};
                const Vec3 t1 = normalize(cross(up, d)), t2 = cross(d, t1);
                const float h = 1e-3f;
                const Vec3 pu = surf(normalize(d + t1 * h)) - surf(normalize(d - t1 * h));
                const Vec3 pv = surf(normalize(d + t2 * h)) - surf(normalize(d - t2 * h));
                Vec3 n = normalize(cross(pu, pv));
                if (dot(n, d) < 0.0f) n = -n;
                m.positions.push_back(pos);
                m.normals.push_back(n);
                m.texcoords.push_back({u, v});
            }
        }
        for (int i = 0; i < freq; i++) {
            for (int j = 0; j < freq - i; j++) {
                const uint32_t p00 = vid(i, j), p10 = vid(i + 1, j), p01 = vid(i, j + 1);
                m.indices.insert(m.indices.end(), {p00, p10, p01});
                if (j < freq - i - 1) {
                    const uint32_t p11 = vid(i + 1, j + 1);
                    m.indices.insert(m.indices.end(), {p10, p11, p01});
                }
            }
        }
        s.meshes.push_back(std::move(m));
    }
}

// Torus (major R, minor r) with nu x nv quads, optional vertex jitter.
static Mesh make_torus(float R, float r, int nu, int nv, float jitter, HostRng *rng) {
    Mesh m;
    for (int i = 0; i < nu; i++) {
        const float u = 6.2831853f * i / nu;
        for (int j = 0; j < nv; j++) {
            const float v = 6.2831853f * j / nv;
            const Vec3 n{std::cos(u) * std::cos(v), std::sin(v), std::sin(u) * std::cos(v)};
            Vec3 p{std::cos(u) * (R + r * std::cos(v)), r * std::sin(v), std::sin(u) * (R + r * std::cos(v))};
            if (rng && jitter > 0.f) p = p + Vec3{rng->normal(), rng->normal(), rng->normal()} * jitter;
            m.positions.push_back(p);
            m.normals.push_back(n);
            m.texcoords.push_back({(float)i / nu, (float)j / nv});
        }
    }
    for (int i = 0; i < nu; i++) {
        for (int j = 0; j < nv; j++) {
            const uint32_t a = (uint32_t)(i * nv + j), b = (uint32_t)(((i + 1) % nu) * nv + j);
            const uint32_t c = (uint32_t)(((i + 1) % nu) * nv + (j + 1) % nv), d = (uint32_t)(i * nv + (j + 1) % nv);
            m.indices.insert(m.indices.end(), {a, d, c, a, c, b});
        }
    }
    return m;
}

// Bevelled box: a subdivided cube pushed onto a box of half-extent `half`
// with edges rounded by `bevel` (6 * n * n * 2 triangles).
static Mesh make_bevelled_box(Vec3 half, float bevel, int n, float jitter, HostRng *rng) {
    Mesh m;
    const Vec3 inner{half.x - bevel, half.y - bevel, half.z - bevel};
    for (int face = 0; face < 6; face++) {
        const int axis = face / 2;
        const float sgn = (face % 2) ? 1.0f : -1.0f;
        const uint32_t base = (uint32_t)m.positions.size();
        for (int i = 0; i <= n; i++) {
            for (int j = 0; j <= n; j++) {
                const float a = -1.0f + 2.0f * i / n, b = -1.0f + 2.0f * j / n;
                float q[3];
                q[axis] = sgn;
                q[(axis + 1) % 3] = a;
                q[(axis + 2) % 3] = b;
                const Vec3 p{q[0] * half.x, q[1] * half.y, q[2] * half.z};
                const Vec3 in{std::fmin(std::fmax(p.x, -inner.x), inner.x), std::fmin(std::fmax(p.y, -inner.y), inner.y),
                              std::fmin(std::fmax(p.z, -inner.z), inner.z)};
                const Vec3 dn = normalize(p - in);
                Vec3 pos = in + dn * bevel;
                if (rng && jitter > 0.f) pos = pos + Vec3{rng->normal(), rng->normal(), rng->normal()} * jitter;
                m.positions.push_back(pos);
                m.normals.push_back(dn);
                m.texcoords.push_back({(float)i / n, (float)j / n});
            }
        }
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++) {
                const uint32_t a = base + (uint32_t)(i * (n + 1) + j), b = a + 1, c = a + (uint32_t)(n + 1), d = c + 1;
                if (sgn > 0) m.indices.insert(m.indices.end(), {a, c, d, a, d, b});
                else m.indices.insert(m.indices.end(), {a, d, c, a, b, d});
            }
        }
    }
    return m;
}

static void transform_mesh(Mesh &m, float scale, Vec3 axis, float angle, Vec3 offset) {
    for (auto &p : m.positions) p = rotate_axis_angle(axis, angle, p * scale) + offset;
    for (auto &n : m.normals) n = normalize(rotate_axis_angle(axis, angle, n));
}

// 50% Lambert / 35% GGX metal (roughness U(0.05, 0.8)) / 15% glass (IOR U(1.3, 1.7)).
static rt_material random_material(HostRng &rng) {
    const float pick = rng.uniform();
    const Vec3 col{rng.uniform(0.15f, 0.9f), rng.uniform(0.15f, 0.9f), rng.uniform(0.15f, 0.9f)};
    if (pick < 0.5f) return material_lambertian(col);
    if (pick < 0.85f) return material_metal(col, rng.uniform(0.05f, 0.8f));
    return material_dielectric(rng.uniform(1.3f, 1.7f), Vec3{0.95f, 0.95f, 0.95f});
}

// ── the scenes ──

static std::unique_ptr<Scene> scene_cornell_box() {
    // Scene::cornell_box, scene.rs:296-381 (geometry and materials as in the reference)
    auto s = std::make_unique<Scene>();
    s->materials.push_back(material_lambertian({0.73f, 0.73f, 0.73f}));
    s->materials.push_back(material_lambertian({0.65f, 0.05f, 0.05f}));
    s->materials.push_back(material_lambertian({0.12f, 0.45f, 0.15f}));
    s->materials.push_back(material_lambertian({0.73f, 0.73f, 0.73f}));
    s->materials.push_back(material_emissive({25.0f, 20.0f, 14.0f}, 1.0f));
    s->materials.push_back(material_metal({0.95f, 0.64f, 0.54f}, 0.1f));
    s->materials.push_back(material_dielectric(1.5f, {1.0f, 1.0f, 1.0f}));
    s->materials.push_back(material_lambertian({0.1f, 0.1f, 0.7f}));
    add_quad(*s, {-1, -1, 1}, {1, -1, 1}, {1, -1, -1}, {-1, -1, -1}, 0);      // floor
    add_quad(*s, {-1, 2, -1}, {1, 2, -1}, {1, 2, 1}, {-1, 2, 1}, 0);          // ceiling
    add_quad(*s, {-1, -1, -1}, {1, -1, -1}, {1, 2, -1}, {-1, 2, -1}, 3);      // back
    add_quad(*s, {-1, -1, 1}, {-1, -1, -1}, {-1, 2, -1}, {-1, 2, 1}, 1);      // left
    add_quad(*s, {1, -1, -1}, {1, -1, 1}, {1, 2, 1}, {1, 2, -1}, 2);          // right
    add_quad(*s, {-0.3f, 1.99f, -0.3f}, {0.3f, 1.99f, -0.3f}, {0.3f, 1.99f, 0.3f}, {-0.3f, 1.99f, 0.3f}, 4);
    add_cube(*s, {0.3f, -1.0f, -0.4f}, {1.0f, -0.4f, 0.0f}, 0);
    add_cube(*s, {-0.7f, -1.0f, 0.1f}, {-0.4f, -0.8f, 0.8f}, 0);
    add_cube(*s, {-0.2f, -1.0f, -0.7f}, {-0.05f, -0.65f, -0.55f}, 5);
    add_cube(*s, {0.5f, -0.8f, -0.7f}, {0.6f, -0.5f, -0.6f}, 6);
    for (uint32_t t : {10u, 11u}) {
        rt_light_desc l{};
        l.kind = RT_LIGHT_AREA;
        l.triangle_index = t;
        l.color[0] = 25.0f; l.color[1] = 20.0f; l.color[2] = 14.0f;
        l.intensity = 1.0f;
        s->lights.push_back(l);
    }
    return s;
}

// The Cornell box of scene_cornell_box with every texture role in use, for
// parity of the texture branches: a checker albedo on the floor (uv 0..3, so
// the sampler's repeat wraps), a tangent-space normal map on the back wall,
// a metallic-roughness map on a box, an emissive map on the ceiling light (and
// on its area-light samples), and an alpha-cutout quad whose albedo layer is
// smaller than the array (the reference leaves the rest of the layer zero,
// i.e. transparent).
static std::unique_ptr<Scene> scene_textured() {
    auto s = scene_cornell_box();
    auto tex = [](uint32_t w, uint32_t h, auto f) {
        Texture t;
        t.width = w;
        t.height = h;
        t.rgba.resize((size_t)w * h * 4);
        for (uint32_t y = 0; y < h; y++)
            for (uint32_t x = 0; x < w; x++) {
                const std::array<float, 4> c = f(x, y);
                for (int k = 0; k < 4; k++) t.rgba[((size_t)y * w + x) * 4 + k] = c[k];
            }
        return t;
    };
    // albedo layer 0: 8x8-texel checker; layer 1: a 16x16 disc mask (alpha)
    s->textures[0].push_back(tex(64, 64, [](uint32_t x, uint32_t y) {
        const bool a = ((x / 8) + (y / 8)) % 2 == 0;
        return std::array<float, 4>{a ? 0.9f : 0.25f, a ? 0.85f : 0.3f, a ? 0.8f : 0.6f, 1.0f};
    }));
    s->textures[0].push_back(tex(16, 16, [](uint32_t x, uint32_t y) {
        const float dx = (float)x - 7.5f, dy = (float)y - 7.5f;
        const float r = std::sqrt(dx * dx + dy * dy) / 8.0f;
        return std::array<float, 4>{0.9f, 0.4f + 0.3f * r, 0.2f, r < 0.8f ? 1.0f : 0.0f};
    }));
    // metallic-roughness: G = roughness ramp, B = metallic stripes
    s->textures[1].push_back(tex(32, 32, [](uint32_t x, uint32_t y) {
        return std::array<float, 4>{0.0f, (float)x / 31.0f, (y / 4) % 2 ? 1.0f : 0.1f, 1.0f};
    }));
    // emissive: warm centre fading to the rim
    s->textures[2].push_back(tex(16, 16, [](uint32_t x, uint32_t y) {
        const float dx = ((float)x - 7.5f) / 8.0f, dy = ((float)y - 7.5f) / 8.0f;
        const float v = std::max(0.0f, 1.0f - 0.7f * (dx * dx + dy * dy));
        return std::array<float, 4>{v, 0.9f * v, 0.7f * v, 1.0f};
    }));
    // normal map: sinusoidal bumps (tangent space, encoded n * 0.5 + 0.5)
    s->textures[3].push_back(tex(32, 32, [](uint32_t x, uint32_t y) {
        const float a = 0.5f * std::sin((float)x * 0.785398f), b = 0.5f * std::cos((float)y * 0.785398f);
        const float inv = 1.0f / std::sqrt(1.0f + a * a + b * b);
        return std::array<float, 4>{a * inv * 0.5f + 0.5f, b * inv * 0.5f + 0.5f, inv * 0.5f + 0.5f, 1.0f};
    }));
    // floor: checker, uv 0..3
    const uint32_t floor_mat = (uint32_t)s->materials.size();
    rt_material fm = material_lambertian({1.0f, 1.0f, 1.0f});
    fm.tex_albedo = 0;
    s->materials.push_back(fm);
    s->meshes[0].material_id = floor_mat;
    for (auto &uv : s->meshes[0].texcoords) uv = Vec2(uv.x * 3.0f, uv.y * 3.0f);
    // back wall: normal map
    rt_material wm = material_lambertian({0.73f, 0.73f, 0.73f});
    wm.tex_normal = 0;
    s->materials.push_back(wm);
    s->meshes[2].material_id = floor_mat + 1;
    // ceiling light: emissive map (the area lights keep their triangles 10, 11)
    rt_material lm = material_emissive({25.0f, 20.0f, 14.0f}, 1.0f);
    lm.tex_emissive = 0;
    s->materials.push_back(lm);
    s->meshes[5].material_id = floor_mat + 2;
    // tall box: metallic-roughness map
    rt_material mm = material_metal({0.9f, 0.8f, 0.7f}, 0.8f);
    mm.metallic = 1.0f;
    mm.tex_mr = 0;
    s->materials.push_back(mm);
    for (int f = 0; f < 6; f++) s->meshes[6 + f].material_id = floor_mat + 3;
    // alpha-cutout quad standing in the room, albedo layer 1
    rt_material cm = material_lambertian({1.0f, 1.0f, 1.0f});
    cm.tex_albedo = 1;
    cm.flags = RT_MAT_FLAG_ALPHA_CUTOUT;
    s->materials.push_back(cm);
    add_quad(*s, {-0.9f, -1.0f, 0.3f}, {0.1f, -1.0f, 0.3f}, {0.1f, 0.8f, 0.3f}, {-0.9f, 0.8f, 0.3f}, floor_mat + 4);
    return s;
}

static std::unique_ptr<Scene> scene_dark_room() {
    // Scene::dark_room, scene.rs:386-460
    auto s = std::make_unique<Scene>();
    s->materials.push_back(material_metal({0.08f, 0.08f, 0.10f}, 0.07f));
    s->materials.push_back(material_metal({0.16f, 0.16f, 0.20f}, 0.20f));
    s->materials.push_back(material_lambertian({0.03f, 0.03f, 0.04f}));
    s->materials.push_back(material_lambertian({0.04f, 0.04f, 0.05f}));
    s->materials.push_back(material_emissive({40.0f, 36.0f, 28.0f}, 1.0f));
    const float half = 2.0f, tile = 0.5f;
    for (int iz = 0; iz < 8; iz++) {
        for (int ix = 0; ix < 8; ix++) {
            const float x0 = -half + ix * tile, z0 = -half + iz * tile;
            add_quad(*s, {x0, 0, z0 + tile}, {x0 + tile, 0, z0 + tile}, {x0 + tile, 0, z0}, {x0, 0, z0},
                     (ix + iz) % 2 == 0 ? 0u : 1u);
        }
    }
    add_quad(*s, {-half, 0, half}, {-half, 0, -half}, {-half, 3, -half}, {-half, 3, half}, 2);
    add_quad(*s, {half, 0, -half}, {half, 0, half}, {half, 3, half}, {half, 3, -half}, 2);
    add_quad(*s, {-half, 0, -half}, {half, 0, -half}, {half, 3, -half}, {-half, 3, -half}, 2);
    add_quad(*s, {-half, 3, -half}, {half, 3, -half}, {half, 3, half}, {-half, 3, half}, 3);
    const uint32_t light_tri = (uint32_t)s->triangle_count();
    add_quad(*s, {-1.5f, 2.99f, 0.15f}, {1.5f, 2.99f, 0.15f}, {1.5f, 2.99f, -0.15f}, {-1.5f, 2.99f, -0.15f}, 4);
    for (uint32_t k = 0; k < 2; k++) {
        rt_light_desc l{};
        l.kind = RT_LIGHT_AREA;
        l.triangle_index = light_tri + k;
        l.color[0] = 40.0f; l.color[1] = 36.0f; l.color[2] = 28.0f;
        l.intensity = 1.0f;
        s->lights.push_back(l);
    }
    return s;
}

static std::unique_ptr<Scene> scene_smoke() {
    // rustracer-renderer/tests/gpu_layout.rs:66-99
    auto s = std::make_unique<Scene>();
    s->materials.push_back(material_lambertian({0.5f, 0.5f, 0.5f}));
    s->materials.push_back(material_emissive({10.0f, 10.0f, 10.0f}, 1.0f));
    Mesh tri;
    tri.positions = {{-1, -1, 0}, {1, -1, 0}, {0, 1, 0}};
    tri.normals = {{0, 0, 1}, {0, 0, 1}, {0, 0, 1}};
    tri.texcoords = {{0, 0}, {0, 0}, {0, 0}};
    tri.indices = {0, 1, 2};
    tri.material_id = 0;
    s->meshes.push_back(tri);
    Mesh quad;
    quad.positions = {{-1, 0, -1}, {1, 0, -1}, {1, 0, 1}, {-1, 0, 1}};
    quad.normals = {{0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 1, 0}};
    quad.texcoords = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    quad.indices = {0, 1, 2, 0, 2, 3};
    quad.material_id = 1;
    s->meshes.push_back(quad);
    add_environment(*s, {0.55f, 0.6f, 0.75f}, {0.95f, 0.93f, 0.9f}, 0.2f);
    return s;
}

static std::unique_ptr<Scene> scene_cornell36() {
    // Unit Cornell room [0,1]^3 open at z = 1: five Lambert walls (left red
    // 0.63/0.065/0.05, right green 0.14/0.45/0.09, others white 0.73), a
    // ceiling light quad of 0.25 x 0.2 of the ceiling with radiance 15, and a
    // tall and a short box rotated about +Y: 10 + 2 + 24 = 36 triangles.
    auto s = std::make_unique<Scene>();
    const uint32_t white = add_material(*s, material_lambertian({0.73f, 0.73f, 0.73f}));
    const uint32_t red = add_material(*s, material_lambertian({0.63f, 0.065f, 0.05f}));
    const uint32_t green = add_material(*s, material_lambertian({0.14f, 0.45f, 0.09f}));
    add_quad(*s, {0, 0, 1}, {1, 0, 1}, {1, 0, 0}, {0, 0, 0}, white);  // floor, normal +Y
    add_quad(*s, {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}, white);  // ceiling, normal -Y
    add_quad(*s, {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, white);  // back, normal +Z
    add_quad(*s, {0, 0, 1}, {0, 0, 0}, {0, 1, 0}, {0, 1, 1}, red);    // left, normal +X
    add_quad(*s, {1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}, green);  // right, normal -X
    const float y = 0.999f, x0 = 0.375f, x1 = 0.625f, z0 = 0.4f, z1 = 0.6f;
    add_area_light_quad(*s, {x0, y, z0}, {x1, y, z0}, {x1, y, z1}, {x0, y, z1}, {15.0f, 15.0f, 15.0f});
    add_rotated_box(*s, 0.33f, 0.36f, 0.0f, {0.28f, 0.6f, 0.28f}, 0.2618f, white);   // tall, +15 deg
    add_rotated_box(*s, 0.66f, 0.62f, 0.0f, {0.28f, 0.3f, 0.28f}, -0.3142f, white);  // short, -18 deg
    return s;
}

static std::unique_ptr<Scene> scene_photon_plain(uint64_t seed) {
    // Displaced icosphere (geodesic frequency 48 = 46 080 triangles, value-noise
    // amplitude 0.05, GGX metal with roughness U(0.1, 0.6) per face group), glass
    // sphere (frequency 20 = 8 000 triangles, IOR 1.5), Lambert floor, one
    // area-light quad and one point light.
    auto s = std::make_unique<Scene>();
    HostRng rng(seed);
    std::vector<uint32_t> metal_mats;
    for (int f = 0; f < 20; f++)
        metal_mats.push_back(add_material(*s, material_metal({0.9f, 0.72f, 0.5f}, rng.uniform(0.1f, 0.6f))));
    const uint32_t nseed = (uint32_t)rng.next();
    add_geodesic(*s, {-0.55f, 0.55f, 0.0f}, 0.5f, 48,
                 [&](Vec3 d) { return 0.05f * value_noise3(d.x * 4.0f, d.y * 4.0f, d.z * 4.0f, nseed); },
                 metal_mats);
    const uint32_t glass = add_material(*s, material_dielectric(1.5f, {1.0f, 1.0f, 1.0f}));
    add_geodesic(*s, {0.6f, 0.4f, 0.25f}, 0.4f, 20, [](Vec3) { return 0.0f; }, std::vector<uint32_t>{glass});
    const uint32_t floor = add_material(*s, material_lambertian({0.7f, 0.7f, 0.68f}));
    add_quad(*s, {-2, 0, 2}, {2, 0, 2}, {2, 0, -2}, {-2, 0, -2}, floor);
    add_area_light_quad(*s, {-0.4f, 2.2f, -0.4f}, {0.4f, 2.2f, -0.4f}, {0.4f, 2.2f, 0.4f}, {-0.4f, 2.2f, 0.4f},
                        {15.0f, 15.0f, 15.0f});
    add_point_light(*s, {1.2f, 1.5f, 1.0f}, {1.0f, 0.95f, 0.9f}, 6.0f);
    return s;
}

static std::unique_ptr<Scene> scene_multiview(uint64_t seed) {
    // ~109k triangles: 30 jittered tori (48 x 24 quads) and 40 bevelled boxes
    // (n = 9), seeded vertex jitter sigma = 0.002, materials 50/35/15.
    auto s = std::make_unique<Scene>();
    HostRng rng(seed);
    for (int i = 0; i < 70; i++) {
        // draws are named one per statement: the order of argument evaluation is unspecified
        Mesh m;
        if (i < 30) {
            const float R = rng.uniform(0.25f, 0.45f);
            const float r = rng.uniform(0.06f, 0.14f);
            m = make_torus(R, r, 48, 24, 0.002f, &rng);
        } else {
            const Vec3 half{rng.uniform(0.15f, 0.35f), rng.uniform(0.15f, 0.35f), rng.uniform(0.15f, 0.35f)};
            m = make_bevelled_box(half, 0.05f, 9, 0.002f, &rng);
        }
        const Vec3 axis = normalize(Vec3{rng.uniform(-1, 1), rng.uniform(-1, 1) + 0.01f, rng.uniform(-1, 1)});
        const int gx = i % 9, gz = i / 9;
        const Vec3 off{-3.2f + 0.8f * gx + rng.uniform(-0.15f, 0.15f), 0.45f + rng.uniform(0.0f, 0.5f),
                       -3.2f + 0.8f * gz + rng.uniform(-0.15f, 0.15f)};
        transform_mesh(m, 1.0f, axis, rng.uniform(0.0f, 6.2831853f), off);
        m.material_id = add_material(*s, random_material(rng));
        s->meshes.push_back(std::move(m));
    }
    const uint32_t ground = add_material(*s, material_lambertian({0.6f, 0.6f, 0.6f}));
    add_quad(*s, {-5, 0, 5}, {5, 0, 5}, {5, 0, -5}, {-5, 0, -5}, ground);
    add_area_light_quad(*s, {-1.5f, 4.0f, -1.0f}, {-0.5f, 4.0f, -1.0f}, {-0.5f, 4.0f, 0.0f}, {-1.5f, 4.0f, 0.0f},
                        {20.0f, 19.0f, 17.0f});
    add_area_light_quad(*s, {1.0f, 3.5f, 1.0f}, {1.8f, 3.5f, 1.0f}, {1.8f, 3.5f, 1.8f}, {1.0f, 3.5f, 1.8f},
                        {12.0f, 14.0f, 18.0f});
    add_environment(*s, {0.45f, 0.55f, 0.8f}, {0.9f, 0.88f, 0.85f}, 0.6f);
    return s;
}

static float terrain_height(float x, float z, uint32_t seed) {
    return 2.2f * fbm3(x * 0.18f, 0.5f, z * 0.18f, seed, 6);
}

static std::unique_ptr<Scene> scene_fullframe(uint64_t seed) {
    // ~1M triangles: value-noise terrain of 548 x 548 quads (600 608 triangles)
    // plus 200 instances of a 2 000-triangle torus (50 x 20 quads) with seeded
    // transforms and materials, flattened into one triangle list.
    auto s = std::make_unique<Scene>();
    HostRng rng(seed);
    const uint32_t nseed = (uint32_t)rng.next();
    const int N = 548;
    const float ext = 12.0f;
    Mesh terrain;
    terrain.material_id = add_material(*s, material_lambertian({0.55f, 0.5f, 0.42f}));
    terrain.positions.reserve((size_t)(N + 1) * (N + 1));
    for (int i = 0; i <= N; i++) {
        for (int j = 0; j <= N; j++) {
            const float x = -ext + 2.0f * ext * j / N, z = -ext + 2.0f * ext * i / N;
            const float h = terrain_height(x, z, nseed);
            const float e = 1e-3f;
            const float dx = (terrain_height(x + e, z, nseed) - terrain_height(x - e, z, nseed)) / (2 * e);
            const float dz = (terrain_height(x, z + e, nseed) - terrain_height(x, z - e, nseed)) / (2 * e);
            terrain.positions.push_back({x, h, z});
            terrain.normals.push_back(normalize(Vec3{-dx, 1.0f, -dz}));
            terrain.texcoords.push_back({(float)j / N, (float)i / N});
        }
    }
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) {
            const uint32_t a = (uint32_t)(i * (N + 1) + j), b = a + 1, c = a + (uint32_t)(N + 1), d = c + 1;
            terrain.indices.insert(terrain.indices.end(), {a, c, d, a, d, b});
        }
    }
    s->meshes.push_back(std::move(terrain));
    const Mesh proto = make_torus(1.0f, 0.35f, 50, 20, 0.0f, nullptr);
    for (int k = 0; k < 200; k++) {
        Mesh m = proto;
        const float sc = rng.uniform(0.15f, 0.45f);
        const float x = rng.uniform(-ext * 0.85f, ext * 0.85f), z = rng.uniform(-ext * 0.85f, ext * 0.85f);
        const Vec3 axis = normalize(Vec3{rng.uniform(-1, 1), rng.uniform(-1, 1) + 0.01f, rng.uniform(-1, 1)});
        transform_mesh(m, sc, axis, rng.uniform(0.0f, 6.2831853f),
                       {x, terrain_height(x, z, nseed) + sc * 1.2f, z});
        m.material_id = add_material(*s, random_material(rng));
        s->meshes.push_back(std::move(m));
    }
    add_area_light_quad(*s, {-3.0f, 9.0f, -3.0f}, {3.0f, 9.0f, -3.0f}, {3.0f, 9.0f, 3.0f}, {-3.0f, 9.0f, 3.0f},
                        {6.0f, 5.6f, 5.0f});
    add_environment(*s, {0.42f, 0.55f, 0.85f}, {0.92f, 0.9f, 0.86f}, 1.0f);
    return s;
}

static std::unique_ptr<Scene> scene_training(uint64_t seed) {
    // One procedural training scene: 10k-200k triangles of instanced tori,
    // bevelled boxes and geodesic spheres, random materials, 1-3 area lights
    // with radiance log-uniform in [1, 50], on a Lambert ground.
    auto s = std::make_unique<Scene>();
    HostRng rng(seed);
    const float target = rng.log_uniform(1e4f, 2e5f);
    const uint32_t ground = add_material(*s, material_lambertian({rng.uniform(0.3f, 0.8f), rng.uniform(0.3f, 0.8f),
                                                                   rng.uniform(0.3f, 0.8f)}));
    add_quad(*s, {-4, 0, 4}, {4, 0, 4}, {4, 0, -4}, {-4, 0, -4}, ground);
    while ((float)s->triangle_count() < target) {
        const float pick = rng.uniform();
        const uint32_t mat = add_material(*s, random_material(rng));
        const Vec3 off{rng.uniform(-2.5f, 2.5f), rng.uniform(0.2f, 1.2f), rng.uniform(-2.5f, 2.5f)};
        // draws are named one per statement: the order of argument evaluation is unspecified
        if (pick < 0.4f) {
            const float R = rng.uniform(0.2f, 0.5f);
            const float r = rng.uniform(0.05f, 0.15f);
            Mesh m = make_torus(R, r, 48, 24, 0.0f, nullptr);
            const Vec3 axis = normalize(Vec3{rng.uniform(-1, 1), 1.0f, rng.uniform(-1, 1)});
            const float angle = rng.uniform(0.0f, 6.2831853f);
            transform_mesh(m, 1.0f, axis, angle, off);
            m.material_id = mat;
            s->meshes.push_back(std::move(m));
        } else if (pick < 0.75f) {
            const Vec3 half{rng.uniform(0.1f, 0.4f), rng.uniform(0.1f, 0.4f), rng.uniform(0.1f, 0.4f)};
            Mesh m = make_bevelled_box(half, 0.04f, 8, 0.0f, nullptr);
            transform_mesh(m, 1.0f, normalize(Vec3{0.0f, 1.0f, 0.0f}), rng.uniform(0.0f, 6.2831853f), off);
            m.material_id = mat;
            s->meshes.push_back(std::move(m));
        } else {
            add_geodesic(*s, off, rng.uniform(0.15f, 0.45f), 16, [](Vec3) { return 0.0f; },
                         std::vector<uint32_t>{mat});
        }
    }
    const int n_lights = 1 + (int)rng.below(3);
    for (int i = 0; i < n_lights; i++) {
        const float L = rng.log_uniform(1.0f, 50.0f);
        const Vec3 c{rng.uniform(-2, 2), rng.uniform(2.5f, 4.0f), rng.uniform(-2, 2)};
        const float h = rng.uniform(0.2f, 0.6f);
        const Vec3 radiance{L, L * rng.uniform(0.8f, 1.0f), L * rng.uniform(0.7f, 1.0f)};
        add_area_light_quad(*s, {c.x - h, c.y, c.z - h}, {c.x + h, c.y, c.z - h}, {c.x + h, c.y, c.z + h},
                            {c.x - h, c.y, c.z + h}, radiance);
    }
    add_environment(*s, {0.45f, 0.55f, 0.8f}, {0.9f, 0.88f, 0.85f}, rng.uniform(0.0f, 0.5f));
    return s;
}

TextureArray build_texture_array(const std::vector<Texture> &textures) {
    TextureArray a;
    if (textures.empty()) return a;
    a.width = std::max(textures[0].width, 1u);
    a.height = std::max(textures[0].height, 1u);
    a.layers = (uint32_t)textures.size();
    a.texels.assign((size_t)a.width * a.height * a.layers * 4, 0);
    for (uint32_t l = 0; l < a.layers; l++) {
        const Texture &t = textures[l];
        const uint32_t w = std::min(t.width, a.width), h = std::min(t.height, a.height);
        for (uint32_t y = 0; y < h; y++)
            for (uint32_t x = 0; x < w; x++)
                for (int c = 0; c < 4; c++) {
                    const float v = t.rgba[((size_t)y * t.width + x) * 4 + c];
                    // Rust: clamp keeps NaN and `NaN as u8` is 0
                    const float q = (v == v) ? std::min(std::max(v, 0.0f), 1.0f) * 255.0f : 0.0f;
                    a.texels[(((size_t)l * a.height + y) * a.width + x) * 4 + c] = (uint8_t)q;
                }
    }
    return a;
}

std::unique_ptr<Scene> builtin_scene(const std::string &name, uint64_t seed) {
    if (name == "cornell_box") return scene_cornell_box();
    if (name == "textured") return scene_textured();
    if (name == "dark_room") return scene_dark_room();
    if (name == "smoke") return scene_smoke();
    if (name == "cornell36") return scene_cornell36();
    if (name == "photon_plain") return scene_photon_plain(seed);
    if (name == "multiview") return scene_multiview(seed);
    if (name == "fullframe") return scene_fullframe(seed);
    if (name == "training") return scene_training(seed);
    return nullptr;
}

// ───────────────────────────── camera (camera.rs) ─────────────────────────────

rt_camera camera_default() {
    rt_camera c{};
    c.position[2] = 5.0f;
    c.up[1] = 1.0f;
    c.fov_degrees = 60.0f;
    c.near_plane = 0.1f;
    c.aperture = 0.0f;
    c.focus_dist = 5.0f;
    return c;
}

static Vec3 V(const float *p) { return {p[0], p[1], p[2]}; }
static void S(float *p, Vec3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

rt_camera_uniform camera_to_uniform(const rt_camera &c, uint32_t w, uint32_t h) {
    const Vec3 forward = normalize(V(c.look_at) - V(c.position));
    const Vec3 right = normalize(cross(forward, V(c.up)));
    const Vec3 up = cross(right, forward);
    const float aspect = (float)w / (float)h;
    const float fov_scale = std::tan(to_radians(c.fov_degrees) * 0.5f);
    rt_camera_uniform u;
    std::memset(&u, 0, sizeof(u));
    S(u.position, V(c.position));
    S(u.forward, forward);
    S(u.right, right * aspect * fov_scale);
    S(u.up, up * fov_scale);
    u.inv_half_res[0] = 2.0f / (float)w;
    u.inv_half_res[1] = 2.0f / (float)h;
    u.near_plane = c.near_plane;
    u.fov_scale = fov_scale;
    u.aperture = c.aperture;
    u.focus_dist = c.focus_dist;
    return u;
}

rt_camera camera_frame_scene(const Scene &s) {
    Vec3 bmin, bmax;
    if (s.bounds(bmin, bmax)) {
        const Vec3 center = (bmin + bmax) * 0.5f;
        const float diag = std::fmax(length(bmax - bmin), 1e-3f);
        const float fov = to_radians(70.0f);
        const float dist = (diag * 0.5f / std::tan(fov * 0.5f)) * 1.3f;
        const Vec3 dir = normalize(Vec3{0.55f, 0.35f, 0.75f});
        rt_camera c = camera_default();
        S(c.position, center + dir * dist);
        S(c.look_at, center);
        S(c.up, {0.f, 1.f, 0.f});
        c.fov_degrees = 70.0f;
        return c;
    }
    return camera_default();
}

void camera_orbit(rt_camera &c, float yaw, float pitch) {
    const Vec3 forward = normalize(V(c.look_at) - V(c.position));
    const float dist = length(V(c.look_at) - V(c.position));
    const Vec3 new_forward = rotate_axis_angle(V(c.up), yaw, forward);
    const Vec3 right = normalize(cross(new_forward, V(c.up)));
    const Vec3 final_forward = rotate_axis_angle(right, pitch, new_forward);
    S(c.position, V(c.look_at) - final_forward * dist);
}

void camera_zoom(rt_camera &c, float delta) {
    const Vec3 forward = normalize(V(c.look_at) - V(c.position));
    const float dist = length(V(c.look_at) - V(c.position));
    const float nd = std::fmax(dist - delta, 0.01f);
    S(c.position, V(c.look_at) - forward * nd);
}

void project_marker(const rt_camera &c, uint32_t w, uint32_t h, Vec3 pos, float radius, float out[4]) {
    const Vec3 position{c.position[0], c.position[1], c.position[2]};
    const Vec3 look{c.look_at[0], c.look_at[1], c.look_at[2]};
    const Vec3 forward = normalize_or_zero(look - position);
    const Vec3 right = normalize_or_zero(cross(forward, Vec3{c.up[0], c.up[1], c.up[2]}));
    const Vec3 up = cross(right, forward);
    const Vec3 v = pos - position;
    const float z = dot(v, forward);
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    if (z <= std::fmax(c.near_plane, 1e-4f)) return;  // behind the camera: hidden
    const float fov_scale = std::tan(to_radians(c.fov_degrees) * 0.5f);
    const float aspect = (float)w / (float)h;
    const float x_ndc = dot(v, right) / (z * fov_scale * aspect);
    const float y_ndc = dot(v, up) / (z * fov_scale);
    // off screen it stays flagged only within 1.5 NDC; the tonemap clips by pixel anyway
    out[3] = (std::fabs(x_ndc) <= 1.5f && std::fabs(y_ndc) <= 1.5f) ? length(v) : 0.0f;
    out[0] = (x_ndc * 0.5f + 0.5f) * (float)w;
    out[1] = (0.5f - y_ndc * 0.5f) * (float)h;
    out[2] = std::fmax((radius / (z * fov_scale)) * ((float)h * 0.5f), 1.5f);
}

void camera_pan(rt_camera &c, float dx, float dy) {
    const Vec3 forward = normalize(V(c.look_at) - V(c.position));
    const Vec3 right = normalize(cross(forward, V(c.up)));
    const Vec3 up = cross(right, forward);
    const Vec3 off = right * dx + up * dy;
    S(c.position, V(c.position) + off);
    S(c.look_at, V(c.look_at) + off);
}

bool builtin_camera(const std::string &name, uint64_t seed, uint32_t index, rt_camera &out) {
    rt_camera c = camera_default();
    auto set = [&](Vec3 pos, Vec3 look, float fov) {
        S(c.position, pos);
        S(c.look_at, look);
        S(c.up, {0.f, 1.f, 0.f});
        c.fov_degrees = fov;
    };
    if (name == "cornell_box" || name == "textured") set({0.0f, 0.5f, 3.0f}, {0, 0, 0}, 50.0f);  // cornell.rs:32-38
    else if (name == "cornell36") set({0.5f, 0.5f, 2.4f}, {0.5f, 0.5f, 0.0f}, 40.0f);
    else if (name == "photon_plain") set({0.0f, 1.3f, 3.4f}, {0.0f, 0.45f, 0.0f}, 45.0f);
    else if (name == "fullframe") set({-7.5f, 6.0f, 9.5f}, {0.5f, 0.0f, 0.0f}, 55.0f);
    else if (name == "multiview" || name == "training") {
        HostRng rng(seed * 7919u + index * 104729u + 17u);
        const float az = rng.uniform(0.0f, 6.2831853f), el = rng.uniform(0.18f, 0.85f);
        const float dist = name == "multiview" ? rng.uniform(3.0f, 7.0f) : rng.uniform(2.5f, 5.0f);
        const Vec3 look{rng.uniform(-0.8f, 0.8f), rng.uniform(0.2f, 0.7f), rng.uniform(-0.8f, 0.8f)};
        const Vec3 dir{std::cos(az) * std::cos(el), std::sin(el), std::sin(az) * std::cos(el)};
        set(look + dir * dist, look, rng.uniform(35.0f, 65.0f));
    } else {
        auto s = builtin_scene(name, seed);
        if (!s) return false;
        c = camera_frame_scene(*s);
    }
    out = c;
    return true;
}

}  // namespace rt
