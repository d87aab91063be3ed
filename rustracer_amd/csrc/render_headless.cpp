// render_headless — the argument set of rustracer-app/examples/render_headless.rs:5-11,18-66
// on the gfx950 backend:
//
//   render_headless <scene> [frames] [out.png] [denoise] [app_like] [photons] [spp]
//                   [accum] [w] [h] [debug_view] [zoom] [yaw] [pitch] [raw_save]
//
// <scene> names a built-in procedural scene (cornell_box, dark_room, smoke,
// cornell36, photon_plain, multiview, fullframe, training), optionally with
// ":seed" appended; glTF files go through `python -m rustracer_amd.render_headless`
// (same arguments), whose loader is Python. Defaults,
// config and camera follow render_headless.rs:18-114 (photon_count 50 000,
// app_like = spp 4 + accumulation 32 + denoise, accum defaults to frames,
// Camera::frame_scene then zoom/orbit). The result is the accumulation buffer,
// saved as a PNG with the CPU tonemap of examples/common/mod.rs:37-47 or, with
// raw_save, as raw little-endian f32 RGBA in <out>.bin.
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rustracer_hip.h"

namespace {

uint32_t crc32(const uint8_t *d, size_t n, uint32_t c = 0) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t r = i;
            for (int k = 0; k < 8; k++) r = (r & 1) ? 0xEDB88320u ^ (r >> 1) : r >> 1;
            table[i] = r;
        }
        init = true;
    }
    c = ~c;
    for (size_t i = 0; i < n; i++) c = table[(c ^ d[i]) & 0xFF] ^ (c >> 8);
    return ~c;
}

void put32(std::vector<uint8_t> &v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}

void chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data) {
    put32(out, (uint32_t)data.size());
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    put32(out, crc32(td.data(), td.size()));
}

// RGBA8 PNG with uncompressed (stored) deflate blocks.
bool write_png(const char *path, const uint8_t *rgba, uint32_t w, uint32_t h) {
    std::vector<uint8_t> raw;
    raw.reserve((size_t)(w * 4 + 1) * h);
    for (uint32_t y = 0; y < h; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba + (size_t)y * w * 4, rgba + (size_t)(y + 1) * w * 4);
    }
    std::vector<uint8_t> z{0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        z.push_back(pos + len == raw.size() ? 1 : 0);
        z.push_back((uint8_t)(len & 0xFF));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xFF));
        z.push_back((uint8_t)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + (ptrdiff_t)pos, raw.begin() + (ptrdiff_t)(pos + len));
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521;
        b = (b + a) % 65521;
    }
    put32(z, (b << 16) | a);
    std::vector<uint8_t> png{0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, w);
    put32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE *f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    std::fclose(f);
    return ok;
}

bool is_one(int argc, char **argv, int i) { return argc > i && std::strcmp(argv[i], "1") == 0; }

// Rust's `str::parse::<u32>()`: an optional '+', then decimal digits only, no
// whitespace, at most u32::MAX. Anything else fails, and the argument keeps its
// default (render_headless.rs:21-24 `.parse().ok().unwrap_or(default)`).
bool parse_u32(const char *s, uint32_t *out) {
    if (*s == '+') s++;
    if (!*s) return false;
    uint64_t v = 0;
    for (; *s; s++) {
        if (*s < '0' || *s > '9') return false;
        v = v * 10 + (uint64_t)(*s - '0');
        if (v > 0xFFFFFFFFull) return false;
    }
    *out = (uint32_t)v;
    return true;
}

bool ieq(const char *a, const char *b) {
    for (; *a && *b; a++, b++)
        if (std::tolower((unsigned char)*a) != *b) return false;
    return *a == *b;
}

// Rust's `str::parse::<f32>()` grammar: [+-] ( inf | infinity | nan | Number ),
// Number = ( digits | digits '.' digits? | '.' digits ) ( [eE] [+-]? digits )?,
// case-insensitive keywords, no whitespace, no hex. A string of that form is
// then converted by strtof, which rounds to nearest as Rust does.
bool parse_f32(const char *s, float *out) {
    const char *p = s;
    if (*p == '+' || *p == '-') p++;
    if (ieq(p, "inf") || ieq(p, "infinity") || ieq(p, "nan")) {
        *out = std::strtof(s, nullptr);
        return true;
    }
    int int_digits = 0, frac_digits = 0;
    while (*p >= '0' && *p <= '9') p++, int_digits++;
    if (*p == '.') {
        p++;
        while (*p >= '0' && *p <= '9') p++, frac_digits++;
    }
    if (int_digits + frac_digits == 0) return false;
    if (*p == 'e' || *p == 'E') {
        p++;
        if (*p == '+' || *p == '-') p++;
        int exp_digits = 0;
        while (*p >= '0' && *p <= '9') p++, exp_digits++;
        if (exp_digits == 0) return false;
    }
    if (*p) return false;
    *out = std::strtof(s, nullptr);
    return true;
}

int fail(const char *what) {
    std::fprintf(stderr, "render_headless: %s: %s\n", what, rt_last_error());
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: render_headless <scene[:seed]> [frames] [out.png] [denoise] [app_like] [photons] "
                             "[spp] [accum] [w] [h] [debug_view] [zoom] [yaw] [pitch] [raw_save]\n");
        return 2;
    }
    auto arg_u = [&](int i, uint32_t d) {
        uint32_t v;
        return argc > i && parse_u32(argv[i], &v) ? v : d;
    };
    auto arg_f = [&](int i, float d) {
        float v;
        return argc > i && parse_f32(argv[i], &v) ? v : d;
    };
    std::string scene_arg = argv[1];
    uint64_t seed = 1;
    const size_t colon = scene_arg.find(':');
    if (colon != std::string::npos) {
        seed = std::strtoull(scene_arg.c_str() + colon + 1, nullptr, 10);
        scene_arg.resize(colon);
    }
    const uint32_t frames = arg_u(2, 128);
    const std::string out = argc > 3 ? argv[3] : "render_headless.png";
    const bool denoise = is_one(argc, argv, 4);
    const bool app_like = is_one(argc, argv, 5);
    const bool photons = is_one(argc, argv, 6);
    const uint32_t spp = arg_u(7, app_like ? 4 : 1);
    const uint32_t accum = arg_u(8, app_like ? 32 : frames);
    const uint32_t w = arg_u(9, 800), h = arg_u(10, 600);
    const uint32_t debug_view = arg_u(11, 0);
    const float zoom = arg_f(12, 0.0f), yaw = arg_f(13, 0.0f), pitch = arg_f(14, 0.0f);
    const bool raw_save = is_one(argc, argv, 15);
    if (std::getenv("RT_HEADLESS_PRINT_ARGS")) {
        // the parsed argument set, printed before anything touches the GPU
        std::printf("scene=%s seed=%llu frames=%u out=%s denoise=%d app_like=%d photons=%d spp=%u accum=%u w=%u "
                    "h=%u debug_view=%u zoom=%.9g yaw=%.9g pitch=%.9g raw_save=%d\n",
                    scene_arg.c_str(), (unsigned long long)seed, frames, out.c_str(), denoise, app_like, photons,
                    spp, accum, w, h, debug_view, (double)zoom, (double)yaw, (double)pitch, raw_save);
        return 0;
    }

    const auto ends_with = [&](const char *sfx) {
        const size_t n = std::strlen(sfx);
        return scene_arg.size() >= n && scene_arg.compare(scene_arg.size() - n, n, sfx) == 0;
    };
    if (ends_with(".gltf") || ends_with(".glb")) {
        std::fprintf(stderr, "render_headless: glTF files load through Python: "
                             "python -m rustracer_amd.render_headless %s ...\n", argv[1]);
        return 2;
    }
    rt_scene *scene = nullptr;
    if (rt_scene_builtin(scene_arg.c_str(), seed, &scene) != RT_OK) return fail("scene");
    uint32_t nt = 0, nm = 0, nl = 0, nmesh = 0;
    rt_scene_counts(scene, &nt, &nm, &nl, &nmesh);
    std::printf("scene: %u meshes, %u mats, %u tris, %u lights\n", nmesh, nm, nt, nl);

    rt_config cfg;
    rt_config_default(&cfg);
    cfg.width = w;
    cfg.height = h;
    cfg.spp = spp;
    cfg.accumulate_frames = accum;
    cfg.enable_denoise = (denoise || app_like) ? 1u : 0u;
    cfg.enable_photon_map = photons ? 1u : 0u;
    cfg.photon_count = 50000;
    cfg.debug_view = debug_view;
    rt_renderer *r = nullptr;
    if (rt_renderer_create(&cfg, scene, 0, &r) != RT_OK) return fail("renderer");
    if (photons && rt_rebuild_photons(r) != RT_OK) return fail("rebuild_photons");
    rt_camera cam;
    rt_camera_frame_scene(scene, &cam);
    if (zoom != 0.0f) rt_camera_zoom(&cam, zoom);
    if (yaw != 0.0f || pitch != 0.0f) rt_camera_orbit(&cam, yaw, pitch);
    rt_update_camera(r, &cam);
    if (rt_render_frames(r, frames) != RT_OK) return fail("render");
    if (rt_synchronize(r) != RT_OK) return fail("synchronize");
    std::vector<float> acc((size_t)w * h * 4);
    if (rt_read_accumulation(r, acc.data(), acc.size()) != RT_OK) return fail("read_accumulation");
    if (raw_save) {
        const std::string path = out + ".bin";
        FILE *f = std::fopen(path.c_str(), "wb");
        if (!f || std::fwrite(acc.data(), 4, acc.size(), f) != acc.size()) return fail("write raw");
        std::fclose(f);
        std::printf("saved raw %s\n", path.c_str());
    } else {
        std::vector<uint8_t> px((size_t)w * h * 4);
        if (rt_tonemap_cpu(acc.data(), (size_t)w * h, px.data()) != RT_OK) return fail("tonemap");
        if (!write_png(out.c_str(), px.data(), w, h)) return fail("write png");
        std::printf("saved %s\n", out.c_str());
    }
    rt_renderer_destroy(r);
    rt_scene_free(scene);
    return 0;
}
