// descriptor.cpp — scene descriptor files (scene.toml) applied to a loaded
// scene: environment override, extra lights, material overrides
// (rustracer-core/src/config.rs:15-172). Host-only.
//
// The reference deserializes the file with serde + the toml crate. This file
// reads the part of TOML that descriptors use: comments, [table] and
// [[array of tables]] headers, bare or quoted keys, and values that are
// numbers, strings, booleans or (possibly multi-line) arrays of those. Field
// types follow serde's rules: f32 fields take integers or floats (floats are
// read as f64 and rounded to f32, as serde does), integer fields take integers
// only, fields this descriptor does not know are ignored, and a wrong type or a
// malformed line is an error naming the line.
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <optional>
#include <string>
#include <vector>

#include "capi_util.h"
#include "scene.h"

namespace rt {
namespace {

struct Value {
    enum Kind { INT, FLOAT, STR, BOOL, ARR } kind = INT;
    int64_t i = 0;
    double f = 0.0;
    std::string s;
    bool b = false;
    std::vector<Value> arr;
};
using Table = std::map<std::string, Value>;

struct Document {
    Table root;
    std::map<std::string, Table> tables;
    std::map<std::string, std::vector<Table>> arrays;
};

[[noreturn]] void fail(int line, const std::string &msg) {
    throw ApiError(RT_ERR_INVALID, "scene descriptor line " + std::to_string(line) + ": " + msg);
}

struct Parser {
    const std::string &text;
    size_t pos = 0;
    int line = 1;

    explicit Parser(const std::string &t) : text(t) {}

    bool eof() const { return pos >= text.size(); }
    char peek() const { return eof() ? '\0' : text[pos]; }
    void skip_ws() {  // spaces and tabs
        while (!eof() && (text[pos] == ' ' || text[pos] == '\t')) pos++;
    }
    void skip_comment() {
        if (peek() == '#')
            while (!eof() && text[pos] != '\n') pos++;
    }
    // whitespace, newlines and comments (inside arrays)
    void skip_all() {
        for (;;) {
            skip_ws();
            skip_comment();
            if (peek() == '\r' || peek() == '\n') {
                if (peek() == '\n') line++;
                pos++;
                continue;
            }
            break;
        }
    }
    void end_of_line() {
        skip_ws();
        skip_comment();
        if (peek() == '\r') pos++;
        if (eof()) return;
        if (peek() != '\n') fail(line, "unexpected text after the value");
        pos++;
        line++;
    }

    std::string key() {
        skip_ws();
        std::string k;
        if (peek() == '"') {
            k = string_value();
        } else {
            while (!eof()) {
                const char c = text[pos];
                if (std::isalnum((unsigned char)c) || c == '_' || c == '-') {
                    k += c;
                    pos++;
                } else {
                    break;
                }
            }
        }
        if (k.empty()) fail(line, "expected a key");
        skip_ws();
        if (peek() == '.') fail(line, "dotted keys are not supported in scene descriptors");
        return k;
    }

    std::string string_value() {
        pos++;  // opening quote
        std::string s;
        for (;;) {
            if (eof() || peek() == '\n') fail(line, "unterminated string");
            const char c = text[pos++];
            if (c == '"') break;
            if (c == '\\') {
                if (eof()) fail(line, "unterminated string");
                const char e = text[pos++];
                switch (e) {
                    case 'n': s += '\n'; break;
                    case 't': s += '\t'; break;
                    case 'r': s += '\r'; break;
                    case '"': s += '"'; break;
                    case '\\': s += '\\'; break;
                    default: fail(line, std::string("unsupported escape \\") + e);
                }
            } else {
                s += c;
            }
        }
        return s;
    }

    Value value() {
        skip_ws();
        Value v;
        const char c = peek();
        if (c == '"') {
            v.kind = Value::STR;
            v.s = string_value();
            return v;
        }
        if (c == '[') {
            v.kind = Value::ARR;
            pos++;
            for (;;) {
                skip_all();
                if (peek() == ']') {
                    pos++;
                    break;
                }
                v.arr.push_back(value());
                skip_all();
                if (peek() == ',') {
                    pos++;
                    continue;
                }
                if (peek() == ']') {
                    pos++;
                    break;
                }
                fail(line, "expected ',' or ']' in an array");
            }
            return v;
        }
        // bare token: boolean or number
        size_t end = pos;
        while (end < text.size() && !std::strchr(" \t\r\n,]#", text[end])) end++;
        std::string tok = text.substr(pos, end - pos);
        pos = end;
        if (tok == "true" || tok == "false") {
            v.kind = Value::BOOL;
            v.b = tok == "true";
            return v;
        }
        std::string digits;
        for (char ch : tok)
            if (ch != '_') digits += ch;
        if (digits.empty()) fail(line, "expected a value");
        const bool is_float = digits.find_first_of(".eE") != std::string::npos || digits == "inf" ||
                              digits == "+inf" || digits == "-inf" || digits == "nan" || digits == "+nan" ||
                              digits == "-nan";
        char *stop = nullptr;
        errno = 0;
        if (is_float) {
            v.kind = Value::FLOAT;
            if (digits.find("inf") != std::string::npos) {
                v.f = digits[0] == '-' ? -INFINITY : INFINITY;
            } else if (digits.find("nan") != std::string::npos) {
                v.f = NAN;
            } else {
                v.f = std::strtod(digits.c_str(), &stop);
                if (*stop != '\0') fail(line, "malformed number '" + tok + "'");
            }
        } else {
            v.kind = Value::INT;
            v.i = std::strtoll(digits.c_str(), &stop, 10);
            if (*stop != '\0' || errno == ERANGE) fail(line, "malformed value '" + tok + "'");
        }
        return v;
    }

    Document parse() {
        Document d;
        Table *cur = &d.root;
        while (!eof()) {
            skip_ws();
            skip_comment();
            if (peek() == '\r') pos++;
            if (eof()) break;
            if (peek() == '\n') {
                pos++;
                line++;
                continue;
            }
            if (peek() == '[') {
                const bool arr = pos + 1 < text.size() && text[pos + 1] == '[';
                pos += arr ? 2 : 1;
                const std::string name = key();
                if (peek() != ']') fail(line, "expected ']'");
                pos++;
                if (arr) {
                    if (peek() != ']') fail(line, "expected ']]'");
                    pos++;
                    if (d.tables.count(name)) fail(line, "'" + name + "' is already a table");
                    d.arrays[name].emplace_back();
                    cur = &d.arrays[name].back();
                } else {
                    if (d.tables.count(name) || d.arrays.count(name))
                        fail(line, "table '" + name + "' defined twice");
                    cur = &d.tables[name];
                }
                end_of_line();
                continue;
            }
            const std::string k = key();
            if (peek() != '=') fail(line, "expected '=' after key '" + k + "'");
            pos++;
            const int key_line = line;
            Value v = value();
            if (cur->count(k)) fail(key_line, "duplicate key '" + k + "'");
            (*cur)[k] = std::move(v);
            end_of_line();
        }
        return d;
    }
};

// serde's field conversions, with the line-free message of a type error.
float as_f32(const Value &v, const char *field) {
    if (v.kind == Value::FLOAT) return (float)v.f;
    if (v.kind == Value::INT) return (float)v.i;
    throw ApiError(RT_ERR_INVALID, std::string("scene descriptor: '") + field + "' must be a number");
}
uint64_t as_uint(const Value &v, const char *field, uint64_t max) {
    if (v.kind != Value::INT || v.i < 0 || (uint64_t)v.i > max)
        throw ApiError(RT_ERR_INVALID, std::string("scene descriptor: '") + field + "' must be an unsigned integer");
    return (uint64_t)v.i;
}
std::string as_str(const Value &v, const char *field) {
    if (v.kind != Value::STR)
        throw ApiError(RT_ERR_INVALID, std::string("scene descriptor: '") + field + "' must be a string");
    return v.s;
}
Vec3 as_vec3(const Value &v, const char *field) {
    if (v.kind != Value::ARR || v.arr.size() != 3)
        throw ApiError(RT_ERR_INVALID, std::string("scene descriptor: '") + field + "' must be an array of 3 numbers");
    return Vec3{as_f32(v.arr[0], field), as_f32(v.arr[1], field), as_f32(v.arr[2], field)};
}
const Value *get(const Table &t, const char *k) {
    const auto it = t.find(k);
    return it == t.end() ? nullptr : &it->second;
}
std::optional<float> opt_f32(const Table &t, const char *k) {
    const Value *v = get(t, k);
    return v ? std::optional<float>(as_f32(*v, k)) : std::nullopt;
}
std::optional<Vec3> opt_vec3(const Table &t, const char *k) {
    const Value *v = get(t, k);
    return v ? std::optional<Vec3>(as_vec3(*v, k)) : std::nullopt;
}

// The deserialized SceneDescriptor (config.rs:17-60).
struct MaterialOverride {
    std::optional<uint64_t> index;
    std::optional<Vec3> albedo, emissive;
    std::optional<float> roughness, metallic, ior;
    std::optional<std::string> kind;
};

void set3(float *dst, Vec3 v) {
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
}

// config.rs:127-172
void apply_override(rt_material &m, const MaterialOverride &ov) {
    if (ov.kind) {
        if (*ov.kind == "lambertian") m.kind = RT_MAT_LAMBERTIAN;
        else if (*ov.kind == "metal") m.kind = RT_MAT_METAL;
        else if (*ov.kind == "dielectric") m.kind = RT_MAT_DIELECTRIC;
        else if (*ov.kind == "emissive") m.kind = RT_MAT_EMISSIVE;
        // any other kind: left as it is (the reference warns)
    }
    if (m.kind == RT_MAT_METAL && !ov.metallic) m.metallic = 1.0f;
    if (m.kind == RT_MAT_DIELECTRIC) {
        if (!ov.metallic) m.metallic = 0.0f;
        if (!ov.roughness) m.roughness = 0.0f;
        if (!ov.ior) m.ior = 1.5f;
    }
    if (ov.albedo) set3(m.albedo, *ov.albedo);
    if (ov.roughness) m.roughness = *ov.roughness;
    if (ov.metallic) m.metallic = *ov.metallic;
    if (ov.emissive) set3(m.emissive, *ov.emissive);
    if (ov.ior) m.ior = *ov.ior;
    if (m.kind == RT_MAT_EMISSIVE && !ov.emissive) {
        m.emissive[0] = m.albedo[0];
        m.emissive[1] = m.albedo[1];
        m.emissive[2] = m.albedo[2];
    }
}

}  // namespace

// config.rs:62-125 (apply_descriptor) after config.rs:174 (parse_descriptor).
// Every field is read and type-checked before the scene changes, so a bad
// file leaves the scene as it was.
void apply_descriptor(Scene &scene, const std::string &toml) {
    Parser p(toml);
    const Document d = p.parse();
    if (!d.root.empty()) {
        // serde ignores unknown top-level keys; a known key that is not a table is a type error
        for (const char *k : {"environment", "lights", "materials"})
            if (d.root.count(k)) throw ApiError(RT_ERR_INVALID, std::string("scene descriptor: '") + k + "' must be a table");
    }
    if (d.arrays.count("environment")) throw ApiError(RT_ERR_INVALID, "scene descriptor: 'environment' must be a table");
    if (d.tables.count("lights") || d.tables.count("materials"))
        throw ApiError(RT_ERR_INVALID, "scene descriptor: 'lights' and 'materials' are arrays of tables");

    // environment (config.rs:27-37, 64-78)
    std::optional<rt_light_desc> env;
    if (const auto it = d.tables.find("environment"); it != d.tables.end()) {
        const Table &t = it->second;
        rt_light_desc l;
        std::memset(&l, 0, sizeof(l));
        l.kind = RT_LIGHT_ENVIRONMENT;
        l.intensity = opt_f32(t, "intensity").value_or(1.0f);
        set3(l.vec_a, opt_vec3(t, "sky_zenith").value_or(Vec3{0.55f, 0.6f, 0.75f}));
        set3(l.vec_b, opt_vec3(t, "sky_horizon").value_or(Vec3{0.95f, 0.93f, 0.9f}));
        env = l;
    }

    // lights (config.rs:39-48, 80-106)
    std::vector<rt_light_desc> extra;
    if (const auto it = d.arrays.find("lights"); it != d.arrays.end()) {
        for (const Table &t : it->second) {
            const Value *kv = get(t, "type");
            if (!kv) throw ApiError(RT_ERR_INVALID, "scene descriptor: a light has no 'type'");
            const std::string kind = as_str(*kv, "type");
            const Value *iv = get(t, "intensity");
            if (!iv) throw ApiError(RT_ERR_INVALID, "scene descriptor: a light has no 'intensity'");
            rt_light_desc l;
            std::memset(&l, 0, sizeof(l));
            l.intensity = as_f32(*iv, "intensity");
            set3(l.color, opt_vec3(t, "color").value_or(Vec3{1.0f, 1.0f, 1.0f}));
            const std::optional<Vec3> position = opt_vec3(t, "position");
            const std::optional<Vec3> direction = opt_vec3(t, "direction");
            const Value *tv = get(t, "triangle_index");
            const uint32_t tri = tv ? (uint32_t)as_uint(*tv, "triangle_index", 0xFFFFFFFFull) : 0u;
            if (kind == "point") {
                l.kind = RT_LIGHT_POINT;
                set3(l.vec_a, position.value_or(Vec3{0.0f, 0.0f, 0.0f}));
            } else if (kind == "directional") {
                l.kind = RT_LIGHT_DIRECTIONAL;
                set3(l.vec_a, normalize(direction.value_or(Vec3{0.0f, -1.0f, 0.0f})));
            } else if (kind == "area") {
                l.kind = RT_LIGHT_AREA;
                l.triangle_index = tri;
            } else {
                continue;  // unknown kind: skipped (the reference warns)
            }
            extra.push_back(l);
        }
    }

    // materials (config.rs:50-59, 108-123)
    std::vector<MaterialOverride> overrides;
    if (const auto it = d.arrays.find("materials"); it != d.arrays.end()) {
        for (const Table &t : it->second) {
            MaterialOverride ov;
            if (const Value *v = get(t, "index")) ov.index = as_uint(*v, "index", ~0ull);
            ov.albedo = opt_vec3(t, "albedo");
            ov.emissive = opt_vec3(t, "emissive");
            ov.roughness = opt_f32(t, "roughness");
            ov.metallic = opt_f32(t, "metallic");
            ov.ior = opt_f32(t, "ior");
            if (const Value *v = get(t, "kind")) ov.kind = as_str(*v, "kind");
            overrides.push_back(ov);
        }
    }

    if (env) {
        // replace any environment light (the loader's default) instead of adding a second
        std::vector<rt_light_desc> kept;
        for (const rt_light_desc &l : scene.lights)
            if (l.kind != RT_LIGHT_ENVIRONMENT) kept.push_back(l);
        kept.push_back(*env);
        scene.lights = kept;
    }
    for (const rt_light_desc &l : extra) scene.lights.push_back(l);
    for (const MaterialOverride &ov : overrides) {
        if (ov.index) {
            if (*ov.index < scene.materials.size()) apply_override(scene.materials[*ov.index], ov);
            // out of range: skipped (the reference warns)
        } else {
            for (rt_material &m : scene.materials) apply_override(m, ov);
        }
    }
}

}  // namespace rt
