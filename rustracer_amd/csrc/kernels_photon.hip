// kernels_photon.hip — photon tracing (rustracer-renderer/shaders/photon_trace.wgsl)
// and the photon-grid build (photon_count.wgsl, the host scan of
// rustracer-renderer/src/lib.rs:987-1007, photon_scatter.wgsl) for gfx950.
//
// The reference counts with global atomics, scans on the CPU after a read-back
// and scatters with an atomic cursor, so photons inside a cell land in arrival
// order. Here the whole build stays on the GPU and is stable (photons in a
// cell keep ascending slot order):
//   1. keys:    cell key per slot (invalid -> num_cells), written by the photon
//               trace as it stores each record (or computed from an uploaded
//               photon array), and the count of valid slots per 2048-slot tile
//   2. compact: the valid (key, slot) pairs in slot order
//   3. sort:    LSD radix sort of those pairs, 8-bit digits; each pass builds
//               per-workgroup digit histograms in LDS, scans them, and ranks
//               equal digits inside a wavefront with ballots
//   4. meta:    grid_meta [start, count] per cell by binary search of the keys
//   5. gather:  sorted_photons[i] = photons[slot of i]
#include <mutex>

#include "device_common.h"
#include "kernels.h"

namespace rtd {

// ─────────────────────────── photon tracing ───────────────────────────

struct PhotonArgs {
    TraceScene ts;
    const rt_triangle *tris;
    const rt_material *mats;
    const rt_light *lights;
    uint32_t n_lights;
    uint32_t n_tris;
    rt_grid_params gp;
    uint32_t n_photons;
    uint32_t max_depth;
    rt_photon *photons;
    uint32_t *keys;  // cell key per slot, or null
    rt_tex_array tex[RT_TEX_ROLES];
};

// photon_count.wgsl:18-33: the u32 cell of a stored photon, or num_cells when
// the slot counts as empty (power.r <= 0) or the cell is outside the grid.
// u32(f32) saturates, so end-of-axis photons alias into the next row's x = 0.
__device__ __forceinline__ uint32_t photon_cell_key(float x, float y, float z, float power_r,
                                                    const rt_grid_params &gp) {
    if (!(power_r > 0.0f)) return gp.num_cells;
    const uint32_t gx = wgsl_u32((x - gp.room_min_x) / gp.cell_size);
    const uint32_t gy = wgsl_u32((y - gp.room_min_y) / gp.cell_size);
    const uint32_t gz = wgsl_u32((z - gp.room_min_z) / gp.cell_size);
    const uint32_t cell = gx + gy * gp.grid_x + gz * gp.grid_x * gp.grid_y;
    return cell < gp.num_cells ? cell : gp.num_cells;
}

// photon_trace.wgsl:91-97
__device__ __forceinline__ F3 bounce_dir(F3 rd, F3 n, float rough, Rng &rng) {
    const F3 mirror = reflect(rd, n);
    const float blur = wgsl_smoothstep(0.0f, 1.0f, rough);
    if (blur <= 0.001f) return mirror;
    const F3 diff = sample_cosine_hemisphere(rng);
    return normalize(mix(mirror, to_world(diff, n), blur));
}

template <int MODE>
__global__ __launch_bounds__(64) void photon_trace_kernel(PhotonArgs A) {
    __shared__ uint32_t stack_rows[kLdsStack * 64];
    const uint32_t photon_index = blockIdx.x * 64u + threadIdx.x;
    if (photon_index >= A.n_photons) return;
    uint32_t spill_words[kSpillAlloc];
    LaneStack stk;
    stk.init(stack_rows, spill_words);
    TraceCount cnt{nullptr, 0u};
    const uint32_t num_lights = A.n_lights;
    Rng rng = rng_init(photon_index * 1103515245u + 12345u);
    uint32_t li = wgsl_u32(rng_f32(rng) * (float)num_lights);
    li = li < num_lights ? li : num_lights - 1u;
    const rt_light &light = A.lights[li];
    const F3 ldata = ld3(light.data), lcolor = ld3(light.color);
    F3 origin = ldata, dir;
    F3 power;
    // photon_trace.wgsl:121-175
    if (light.kind == 0u) {
        float u1, u2;
        for (;;) {
            u1 = rng_f32(rng) * 2.0f - 1.0f;
            u2 = rng_f32(rng) * 2.0f - 1.0f;
            if (u1 * u1 + u2 * u2 < 1.0f) break;
        }
        const float z = 1.0f - 2.0f * (u1 * u1 + u2 * u2);
        const float r = wgsl_sqrt(1.0f - z * z);
        dir = normalize(f3(r * 2.0f * u1, r * 2.0f * u2, z));
        power = lcolor * light.intensity;
    } else if (light.kind == 1u) {
        const float disk_u = (rng_f32(rng) - 0.5f) * 10.0f;
        const float disk_v = (rng_f32(rng) - 0.5f) * 10.0f;
        const F3 up = (__builtin_fabsf(ldata.y) < 0.999f) ? f3(0.0f, 1.0f, 0.0f) : f3(1.0f, 0.0f, 0.0f);
        const F3 tangent = normalize(cross(up, ldata));
        const F3 bitangent = cross(ldata, tangent);
        origin = disk_u * tangent + disk_v * bitangent - ldata * 100.0f;
        dir = ldata;
        power = lcolor * light.intensity;
    } else if (light.kind == 2u) {
        uint32_t ti = wgsl_u32(light.data[0]);
        ti = ti < A.n_tris ? ti : A.n_tris - 1u;
        const rt_triangle &tri = A.tris[ti];
        const float u1 = rng_f32(rng);
        const float u2 = rng_f32(rng);
        const float su = wgsl_sqrt(u1);
        const float u = 1.0f - su;
        const float v = u2 * su;
        const float w = 1.0f - u - v;
        origin = ld3(tri.v0) * u + ld3(tri.v1) * v + ld3(tri.v2) * w;
        const F3 n = get_triangle_normal(tri);
        const F3 local_dir = sample_cosine_hemisphere(rng);
        dir = to_world(local_dir, n);
        power = lcolor * light.intensity;
    } else if (light.kind == 3u) {
        const F3 local_dir = sample_cosine_hemisphere(rng);
        const F3 d = to_world(local_dir, f3(0.0f, 1.0f, 0.0f));
        const rt_grid_params &gp = A.gp;
        const F3 span = f3((float)gp.grid_x, (float)gp.grid_y, (float)gp.grid_z) * gp.cell_size;
        const F3 center = f3(gp.room_min_x, gp.room_min_y, gp.room_min_z) + span * 0.5f;
        const float R = wmax(wmax(span.x, span.y), span.z) * 1.2f + 0.05f;
        origin = center + d * R;
        dir = -d;
        const float t = wclamp(d.y * 0.5f + 0.5f, 0.0f, 1.0f);
        power = mix(lcolor, ldata, wgsl_pow(t, 0.5f)) * light.intensity;
    } else {
        return;
    }

    // photon_trace.wgsl:178-259
    for (uint32_t bounce = 0u; bounce < A.max_depth; bounce++) {
        const Hit hit = trace_mode<MODE, false, false>(A.ts, origin, dir, kInf, stk, cnt);
        if (hit.t >= kInf) return;
        const rt_triangle &tri = A.tris[hit.prim];
        const rt_material &mat = A.mats[tri.material_id];
        const F3 hit_pos = origin + dir * hit.t;
        const Bary bb = barycentric(tri, origin, dir, hit.t);
        const F3 normal = get_triangle_normal(tri);
        const bool front_face = dot(dir, normal) < 0.0f;
        const F3 sn_geom = smooth_normal(tri, bb, normal);
        const F3 face_normal = front_face ? sn_geom : -sn_geom;
        F3 shn = face_normal;
        if (mat.tex_normal != RT_TEX_NONE) {  // photon_trace.wgsl:193-198
            const F3 nm = rgb(tex_sample(A.tex[RT_TEX_ROLE_NORMAL], bb.uv, mat.tex_normal)) * 2.0f - fs(1.0f);
            const TanFrame frm = tangent_frame(tri, face_normal);
            shn = normalize(frm.t * nm.x + frm.b * nm.y + face_normal * nm.z);
        }
        // photon_trace.wgsl:200-206: pass through a transparent texel of a cutout material
        if ((mat.flags & RT_MAT_FLAG_ALPHA_CUTOUT) != 0u && mat.tex_albedo != RT_TEX_NONE) {
            if (tex_sample(A.tex[RT_TEX_ROLE_ALBEDO], bb.uv, mat.tex_albedo).w < 0.5f) {
                origin = hit_pos + dir * 0.002f;
                continue;
            }
        }
        if (mat.tex_emissive != RT_TEX_NONE || length(ld3(mat.emissive)) > 1e-4f || mat.kind == 3u) return;
        if (mat.kind == 2u) {
            const F3 f0 = fs(wgsl_pow((mat.ior - 1.0f) / (mat.ior + 1.0f), 2.0f));
            const float cos_theta = __builtin_fabsf(dot(dir, face_normal));
            const float fr = fresnel_schlick(cos_theta, f0).x;
            if (rng_f32(rng) < fr) {
                dir = reflect(dir, face_normal);
                origin = hit_pos + face_normal * 0.001f;
            } else {
                const float eta = front_face ? 1.0f / mat.ior : mat.ior;
                dir = refract(dir, face_normal, eta);
                origin = hit_pos - face_normal * 0.001f;
            }
            continue;
        }
        const size_t slot = (size_t)photon_index * A.max_depth + bounce;
        float4 *rec = reinterpret_cast<float4 *>(A.photons + slot);
        rec[0] = make_float4(hit_pos.x, hit_pos.y, hit_pos.z, 0.0f);
        rec[1] = make_float4(power.x, power.y, power.z, 0.0f);
        rec[2] = make_float4(-dir.x, -dir.y, -dir.z, 0.0f);
        rec[3] = make_float4(shn.x, shn.y, shn.z, 0.0f);
        // the grid key of the record just stored, from the same floats
        if (A.keys) A.keys[slot] = photon_cell_key(hit_pos.x, hit_pos.y, hit_pos.z, power.x, A.gp);

        // photon_trace.wgsl:84-89 surface_albedo: the texture alone when present
        const F3 albedo = mat.tex_albedo != RT_TEX_NONE ? rgb(tex_sample(A.tex[RT_TEX_ROLE_ALBEDO], bb.uv, mat.tex_albedo))
                                                        : f3(mat.albedo[0], mat.albedo[1], mat.albedo[2]);
        const F3 f0 = mix(fs(0.04f), albedo, fs(mat.metallic));
        const float ct = __builtin_fabsf(dot(dir, shn));
        const F3 fr = fresnel_schlick(ct, f0);
        const F3 diffuse_weight = (1.0f - mat.metallic) * albedo;
        const float spec_prob = wclamp(vmax3(fr), 0.0f, 0.98f);
        const bool do_spec = rng_f32(rng) < spec_prob;
        if (do_spec) {
            dir = bounce_dir(dir, shn, mat.roughness, rng);
            power = power * f0;
        } else {
            const F3 local_dir = sample_cosine_hemisphere(rng);
            dir = to_world(local_dir, shn);
            power = power * diffuse_weight;
        }
        origin = hit_pos + shn * 0.001f;
        const float p = vmax3(power);
        if (p <= 0.0f || rng_f32(rng) > p) return;
        power = power / p;
    }
}

// ─────────────────────────── grid build ───────────────────────────

constexpr uint32_t kSortThreads = 256;  // 4 wavefronts
constexpr uint32_t kSortItems = 8;      // keys per thread per tile
constexpr uint32_t kSortTile = kSortThreads * kSortItems;
constexpr uint32_t kRadixBits = 8;
constexpr uint32_t kRadix = 1u << kRadixBits;

__device__ __forceinline__ uint32_t wave_reduce_add(uint32_t v) {
#pragma unroll
    for (uint32_t o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Keys of an uploaded photon array (photon_count.wgsl:18-33). One workgroup covers a tile of kSortTile consecutive slots, reads them in
// kSortItems rounds of 256 and writes the key of every slot plus the tile's
// count of valid slots. No atomics: the per-cell counts come from the sorted
// keys afterwards, so hot cells (many photons in one cell) cost nothing extra.
__global__ __launch_bounds__(256) void grid_keys_kernel(const rt_photon *photons, uint32_t n_slots, rt_grid_params gp,
                                                         uint32_t *keys, uint32_t *tile_counts) {
    __shared__ uint32_t lds_wave[4];
    const size_t base = (size_t)blockIdx.x * kSortTile;
    float4 pos[kSortItems];
    float pr[kSortItems];
#pragma unroll
    for (uint32_t r = 0; r < kSortItems; r++) {
        const size_t slot = base + (size_t)r * kSortThreads + threadIdx.x;
        pos[r] = make_float4(0.f, 0.f, 0.f, 0.f);
        pr[r] = 0.0f;
        if (slot < n_slots) {
            pos[r] = *reinterpret_cast<const float4 *>(&photons[slot].position[0]);
            pr[r] = photons[slot].power[0];
        }
    }
    uint32_t valid = 0;
#pragma unroll
    for (uint32_t r = 0; r < kSortItems; r++) {
        const size_t slot = base + (size_t)r * kSortThreads + threadIdx.x;
        const uint32_t key = photon_cell_key(pos[r].x, pos[r].y, pos[r].z, pr[r], gp);
        if (slot < n_slots) {
            keys[slot] = key;
            valid += key < gp.num_cells ? 1u : 0u;
        }
    }
    const uint32_t wave_sum = wave_reduce_add(valid);
    if ((threadIdx.x & 63u) == 0u) lds_wave[threadIdx.x >> 6] = wave_sum;
    __syncthreads();
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = lds_wave[0] + lds_wave[1] + lds_wave[2] + lds_wave[3];
}

// Valid slots per tile of keys written by the photon trace.
__global__ __launch_bounds__(256) void grid_tile_count_kernel(const uint32_t *keys, uint32_t n_slots,
                                                               uint32_t num_cells, uint32_t *tile_counts) {
    __shared__ uint32_t lds_wave[4];
    const size_t base = (size_t)blockIdx.x * kSortTile;
    uint32_t valid = 0;
#pragma unroll
    for (uint32_t r = 0; r < kSortItems; r++) {
        const size_t i = base + (size_t)r * kSortThreads + threadIdx.x;
        if (i < n_slots) valid += keys[i] < num_cells ? 1u : 0u;
    }
    const uint32_t wave_sum = wave_reduce_add(valid);
    if ((threadIdx.x & 63u) == 0u) lds_wave[threadIdx.x >> 6] = wave_sum;
    __syncthreads();
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = lds_wave[0] + lds_wave[1] + lds_wave[2] + lds_wave[3];
}

// Stream compaction that keeps slot order: the valid (key, slot) pairs of each
// tile go to tile_offsets[tile] + their rank among the tile's valid slots.
__global__ __launch_bounds__(256) void grid_compact_kernel(const uint32_t *keys, uint32_t n_slots, uint32_t num_cells,
                                                            const uint32_t *tile_offsets, uint32_t *keys_out,
                                                            uint32_t *vals_out) {
    __shared__ uint32_t wave_tot[kSortItems][4];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t lt_mask = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
    const size_t base = (size_t)blockIdx.x * kSortTile;
    uint32_t key[kSortItems], rank[kSortItems];
#pragma unroll
    for (uint32_t r = 0; r < kSortItems; r++) {
        const size_t i = base + (size_t)r * kSortThreads + threadIdx.x;
        key[r] = i < n_slots ? keys[i] : num_cells;
    }
#pragma unroll
    for (uint32_t r = 0; r < kSortItems; r++) {
        const uint64_t mask = __ballot(key[r] < num_cells);
        rank[r] = (uint32_t)__popcll(mask & lt_mask);
        if (lane == 0u) wave_tot[r][wave] = (uint32_t)__popcll(mask);
    }
    __syncthreads();
    uint32_t run = tile_offsets[blockIdx.x];
#pragma unroll
    for (uint32_t r = 0; r < kSortItems; r++) {
        uint32_t off = run;
        for (uint32_t w = 0; w < wave; w++) off += wave_tot[r][w];
        if (key[r] < num_cells) {
            keys_out[off + rank[r]] = key[r];
            vals_out[off + rank[r]] = (uint32_t)(base + (size_t)r * kSortThreads + threadIdx.x);
        }
        run += wave_tot[r][0] + wave_tot[r][1] + wave_tot[r][2] + wave_tot[r][3];
    }
}

// grid_meta[c] = [start, count] (the reference's exclusive scan of the cell
// counts, lib.rs:987-998) read off the sorted keys: start = number of keys
// below c, count = number of keys equal to c. Cells past num_cells get
// [n_valid, 0], as the scan gives them.
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, len = n;
    while (len > 0) {
        const uint32_t half = len >> 1;
        if (a[lo + half] < v) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

__global__ __launch_bounds__(256) void grid_meta_kernel(const uint32_t *sorted_keys, const uint32_t *n_valid,
                                                         uint32_t meta_cells, uint2 *meta) {
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= meta_cells) return;
    const uint32_t n = *n_valid;
    const uint32_t lo = lower_bound_u32(sorted_keys, n, c);
    const uint32_t hi = lo + lower_bound_u32(sorted_keys + lo, n - lo, c + 1u);
    meta[c] = make_uint2(lo, hi - lo);
}

// Block-wide exclusive scan of one value per thread (256 threads = 4 waves).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds_wave, uint32_t &total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (uint32_t o = 1; o < 64u; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63u) lds_wave[wave] = x;
    __syncthreads();
    uint32_t wave_off = 0, sum = 0;
    for (uint32_t w = 0; w < kSortThreads / 64u; w++) {
        const uint32_t t = lds_wave[w];
        if (w < wave) wave_off += t;
        sum += t;
    }
    __syncthreads();
    total = sum;
    return wave_off + x - v;
}

// Scan phase 1: per-tile sums of `in` (n values).
__global__ __launch_bounds__(256) void scan_reduce_kernel(const uint32_t *in, uint32_t n, uint32_t *tile_sums) {
    __shared__ uint32_t lds_wave[4];
    const size_t base = (size_t)blockIdx.x * kSortTile;
    uint32_t s = 0;
    for (uint32_t k = 0; k < kSortItems; k++) {
        const size_t i = base + (size_t)k * kSortThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t total;
    block_exclusive_scan(s, lds_wave, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Scan phase 2: exclusive scan of the tile sums in one workgroup (any count).
__global__ __launch_bounds__(256) void scan_tiles_kernel(uint32_t *tile_sums, uint32_t n_tiles, uint32_t *grand_total) {
    __shared__ uint32_t lds_wave[4];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n_tiles; base += kSortThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n_tiles ? tile_sums[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, lds_wave, total);
        if (i < n_tiles) tile_sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0 && grand_total) *grand_total = carry;
}

// Scan phase 3: exclusive scan inside each tile plus the tile offset.
// Writes out[i] (or meta[2i] = start, meta[2i+1] = count when meta != null).
__global__ __launch_bounds__(256) void scan_apply_kernel(const uint32_t *in, uint32_t n, const uint32_t *tile_offsets,
                                                          uint32_t *out, uint32_t *meta) {
    __shared__ uint32_t lds_wave[4];
    const size_t base = (size_t)blockIdx.x * kSortTile;
    // each thread owns kSortItems consecutive values
    uint32_t v[kSortItems];
    uint32_t s = 0;
    const size_t my = base + (size_t)threadIdx.x * kSortItems;
#pragma unroll
    for (uint32_t k = 0; k < kSortItems; k++) {
        v[k] = (my + k < n) ? in[my + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = tile_offsets[blockIdx.x] + block_exclusive_scan(s, lds_wave, total);
#pragma unroll
    for (uint32_t k = 0; k < kSortItems; k++) {
        if (my + k < n) {
            if (meta) {
                meta[2 * (my + k)] = run;
                meta[2 * (my + k) + 1] = v[k];
            } else {
                out[my + k] = run;
            }
        }
        run += v[k];
    }
}

// Radix pass, step 1: digit histogram of each tile, stored digit-major
// (hist[d * n_tiles + tile]) so one exclusive scan gives every (digit, tile) base.
__global__ __launch_bounds__(256) void radix_hist_kernel(const uint32_t *keys, const uint32_t *n_dev, uint32_t shift,
                                                         uint32_t n_tiles, uint32_t *hist) {
    __shared__ uint32_t h[kRadix];
    const uint32_t n = *n_dev;
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kSortTile;
    for (uint32_t k = 0; k < kSortItems; k++) {
        const size_t i = base + (size_t)k * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1u)], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// Radix pass, step 2: stable scatter. The tile is consumed in kSortItems rounds
// of 256 consecutive keys; inside a round each wavefront ranks its lanes among
// equal digits with 8 ballots (bit-by-bit match), wavefronts are ordered
// through LDS, and a running per-digit offset carries the order across rounds.
__global__ __launch_bounds__(256) void radix_scatter_kernel(const uint32_t *keys_in, const uint32_t *vals_in,
                                                            const uint32_t *n_dev, uint32_t shift, uint32_t n_tiles,
                                                            const uint32_t *hist_scanned, uint32_t *keys_out,
                                                            uint32_t *vals_out) {
    __shared__ uint32_t digit_base[kRadix];        // global base + items already placed
    __shared__ uint32_t wave_counts[4][kRadix];    // per-round, per-wave digit counts
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t n = *n_dev;
    const size_t base = (size_t)blockIdx.x * kSortTile;
    if (base >= n) return;  // tile past the valid count: nothing to place (uniform per workgroup)
    digit_base[threadIdx.x] = hist_scanned[(size_t)threadIdx.x * n_tiles + blockIdx.x];
    const uint64_t lt_mask = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
    for (uint32_t r = 0; r < kSortItems; r++) {
        for (uint32_t w = 0; w < 4u; w++) wave_counts[w][threadIdx.x] = 0u;
        __syncthreads();
        const size_t i = base + (size_t)r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const uint32_t digit = (key >> shift) & (kRadix - 1u);
        uint64_t match = __ballot(valid);
#pragma unroll
        for (uint32_t bit = 0; bit < kRadixBits; bit++) {
            const uint64_t b = __ballot((digit >> bit) & 1u);
            match &= ((digit >> bit) & 1u) ? b : ~b;
        }
        const uint32_t rank_in_wave = (uint32_t)__popcll(match & lt_mask);
        if (valid && rank_in_wave == 0u) wave_counts[wave][digit] = (uint32_t)__popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t off = digit_base[digit] + rank_in_wave;
            for (uint32_t w = 0; w < wave; w++) off += wave_counts[w][digit];
            keys_out[off] = key;
            vals_out[off] = vals_in[i];
        }
        __syncthreads();
        // advance the running base of every digit by this round's count
        digit_base[threadIdx.x] += wave_counts[0][threadIdx.x] + wave_counts[1][threadIdx.x] +
                                   wave_counts[2][threadIdx.x] + wave_counts[3][threadIdx.x];
        __syncthreads();
    }
}

// photon_scatter.wgsl:35-36 with the sorted order: copy the 64-byte records,
// and write the nine floats the radius gather reads (position, power,
// incident direction) once more as a dense 48-byte record, so one batch of
// candidate loads brings everything the gather needs.
__global__ __launch_bounds__(256) void grid_gather_kernel(const rt_photon *photons, const uint32_t *vals,
                                                          const uint32_t *n_valid, rt_photon *sorted,
                                                          float4 *gather_rec) {
    // 4 lanes per photon, one float4 each; a grid-stride loop over the stored
    // photons only (the launch is sized for the slot count, which is far larger)
    const size_t total = (size_t)*n_valid * 4u;
    for (size_t t = (size_t)blockIdx.x * 256u + threadIdx.x; t < total; t += (size_t)gridDim.x * 256u) {
        const size_t i = t >> 2;
        const uint32_t part = (uint32_t)(t & 3u);
        const float4 *src = reinterpret_cast<const float4 *>(photons + vals[i]);
        float4 *dst = reinterpret_cast<float4 *>(sorted + i);
        const float4 v = src[part];
        dst[part] = v;
        if (!gather_rec) continue;  // uniform: the caller keeps no gather records
        if (part == 0u) {
            gather_rec[3 * i] = make_float4(v.x, v.y, v.z, src[1].x);
        } else if (part == 1u) {
            const float4 inc = src[2];
            gather_rec[3 * i + 1] = make_float4(v.y, v.z, inc.x, inc.y);
        } else if (part == 2u) {
            gather_rec[3 * i + 2] = make_float4(v.z, 0.0f, 0.0f, 0.0f);
        }
    }
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t value, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = value;
}

TraceScene make_trace_scene(const rtk::TraceInputs &t);

}  // namespace rtd

namespace rtk {

hipError_t launch_photon_trace(const PhotonLaunch &p, hipStream_t s) {
    if (p.n_photons == 0) return hipSuccess;
    rtd::PhotonArgs A;
    A.ts = rtd::make_trace_scene(p.trace);
    A.tris = p.tris;
    A.mats = p.mats;
    A.lights = p.lights;
    A.n_lights = p.n_lights;
    A.n_tris = p.n_tris;
    A.gp = p.gp;
    A.n_photons = p.n_photons;
    A.max_depth = p.max_depth;
    A.photons = p.photons;
    A.keys = p.keys;
    for (uint32_t k = 0; k < RT_TEX_ROLES; k++) A.tex[k] = p.tex[k];
    const dim3 grid((p.n_photons + 63u) / 64u), block(64);
    if (p.mode == 0) hipLaunchKernelGGL((rtd::photon_trace_kernel<0>), grid, block, 0, s, A);
    else if (p.mode == 1) hipLaunchKernelGGL((rtd::photon_trace_kernel<1>), grid, block, 0, s, A);
    else if (p.mode == 2) hipLaunchKernelGGL((rtd::photon_trace_kernel<2>), grid, block, 0, s, A);
    else if (p.mode == 4) hipLaunchKernelGGL((rtd::photon_trace_kernel<4>), grid, block, 0, s, A);
    else hipLaunchKernelGGL((rtd::photon_trace_kernel<3>), grid, block, 0, s, A);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(rtd::fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, value, n);
    return hipGetLastError();
}

static constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
static constexpr uint32_t tiles_of(size_t n) { return (uint32_t)((n + rtd::kSortTile - 1) / rtd::kSortTile); }
// Device temporaries of a build over n slots (laid out by launch_grid_build).
static constexpr size_t temp_bytes_for(size_t n) {
    const uint32_t t = tiles_of(n) > 1u ? tiles_of(n) : 1u;
    const size_t hist = (size_t)rtd::kRadix * t;
    const uint32_t th = tiles_of(hist) > 1u ? tiles_of(hist) : 1u;
    return 4 * align256(n * 4) +                      // keys/vals x 2
           align256(hist * 4) +                       // histograms
           align256((size_t)t * 4) +                  // valid slots per tile
           align256((size_t)(t > th ? t : th) * 4) +  // scan tile sums
           256;
}

size_t grid_build_temp_bytes(uint32_t n_slots, uint32_t meta_cells) {
    (void)meta_cells;
    return temp_bytes_for(n_slots);
}

// Exclusive scan of `in[0..n)` into out.
static hipError_t scan_u32(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *tile_sums, hipStream_t s) {
    const uint32_t tiles = std::max(1u, tiles_of(n));
    hipLaunchKernelGGL(rtd::scan_reduce_kernel, dim3(tiles), dim3(256), 0, s, in, n, tile_sums);
    hipLaunchKernelGGL(rtd::scan_tiles_kernel, dim3(1), dim3(256), 0, s, tile_sums, tiles, nullptr);
    hipLaunchKernelGGL(rtd::scan_apply_kernel, dim3(tiles), dim3(256), 0, s, in, n, tile_sums, out, nullptr);
    return hipGetLastError();
}

// photon_count + CPU scan + photon_scatter (photon_count.wgsl, lib.rs:987-998,
// photon_scatter.wgsl) as one device pipeline: keys of every slot, compaction
// of the valid slots in slot order, a stable LSD radix sort of those pairs
// (so photons of one cell stay in ascending slot order), grid_meta from the
// sorted keys, then the record copy in sorted order.
hipError_t launch_grid_build(const GridBuildLaunch &g, hipStream_t s) {
    const uint32_t n = g.n_slots;
    const uint32_t tiles = std::max(1u, tiles_of(n));
    const size_t hist_n = (size_t)rtd::kRadix * tiles;
    char *p = reinterpret_cast<char *>(g.temp);
    uint32_t *keys_a = reinterpret_cast<uint32_t *>(p); p += align256((size_t)n * 4);
    uint32_t *vals_a = reinterpret_cast<uint32_t *>(p); p += align256((size_t)n * 4);
    uint32_t *keys_b = reinterpret_cast<uint32_t *>(p); p += align256((size_t)n * 4);
    uint32_t *vals_b = reinterpret_cast<uint32_t *>(p); p += align256((size_t)n * 4);
    uint32_t *hist = reinterpret_cast<uint32_t *>(p); p += align256(hist_n * 4);
    uint32_t *tile_counts = reinterpret_cast<uint32_t *>(p); p += align256((size_t)tiles * 4);
    uint32_t *tile_sums = reinterpret_cast<uint32_t *>(p);
    hipError_t e;
    const unsigned meta_blocks = (unsigned)((g.meta_cells + 255u) / 256u);
    if (n == 0) {
        if ((e = launch_fill_u32(g.n_valid_out, 0u, 1, s)) != hipSuccess) return e;
        if (meta_blocks)
            hipLaunchKernelGGL(rtd::grid_meta_kernel, dim3(meta_blocks), dim3(256), 0, s, keys_a, g.n_valid_out,
                               g.meta_cells, reinterpret_cast<uint2 *>(g.grid_meta));
        return hipGetLastError();
    }
    // keys of all slots (from the trace, or computed into keys_b, which is
    // free until the first radix pass) and the valid slots per tile
    const uint32_t *keys = g.keys_in;
    if (keys) {
        hipLaunchKernelGGL(rtd::grid_tile_count_kernel, dim3(tiles), dim3(256), 0, s, keys, n, g.gp.num_cells,
                           tile_counts);
    } else {
        hipLaunchKernelGGL(rtd::grid_keys_kernel, dim3(tiles), dim3(256), 0, s, g.photons, n, g.gp, keys_b,
                           tile_counts);
        keys = keys_b;
    }
    // tile offsets of the compacted list; n_valid = their total
    hipLaunchKernelGGL(rtd::scan_tiles_kernel, dim3(1), dim3(256), 0, s, tile_counts, tiles, g.n_valid_out);
    hipLaunchKernelGGL(rtd::grid_compact_kernel, dim3(tiles), dim3(256), 0, s, keys, n, g.gp.num_cells, tile_counts,
                       keys_a, vals_a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // key bits needed for the valid keys 0..num_cells-1
    uint32_t bits = 0;
    const uint64_t max_key = g.gp.num_cells > 0 ? (uint64_t)g.gp.num_cells - 1 : 0;
    while (bits < 32 && (1ull << bits) <= max_key) bits++;
    const uint32_t passes = std::max(1u, (bits + rtd::kRadixBits - 1) / rtd::kRadixBits);
    uint32_t *kin = keys_a, *vin = vals_a, *kout = keys_b, *vout = vals_b;
    for (uint32_t pass = 0; pass < passes; pass++) {
        const uint32_t shift = pass * rtd::kRadixBits;
        hipLaunchKernelGGL(rtd::radix_hist_kernel, dim3(tiles), dim3(256), 0, s, kin, g.n_valid_out, shift, tiles,
                           hist);
        if ((e = scan_u32(hist, (uint32_t)hist_n, hist, tile_sums, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(rtd::radix_scatter_kernel, dim3(tiles), dim3(256), 0, s, kin, vin, g.n_valid_out, shift,
                           tiles, hist, kout, vout);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (meta_blocks)
        hipLaunchKernelGGL(rtd::grid_meta_kernel, dim3(meta_blocks), dim3(256), 0, s, kin, g.n_valid_out, g.meta_cells,
                           reinterpret_cast<uint2 *>(g.grid_meta));
    const size_t threads = (size_t)n * 4;
    const unsigned gather_blocks = (unsigned)std::min<size_t>((threads + 255) / 256, 2048);
    hipLaunchKernelGGL(rtd::grid_gather_kernel, dim3(gather_blocks), dim3(256), 0, s, g.photons,
                       vin, g.n_valid_out, g.sorted, reinterpret_cast<float4 *>(g.gather_rec));
    return hipGetLastError();
}

// The allocation-free build's workspace: a file-scope device array sized for
// kNoallocMaxSlots slots. Every GPU the library runs on holds its own copy, at
// its own address, until the process ends; calls on one GPU share it (the
// last call's contents stay), so calls that overlap on one GPU race. No part
// of it needs clearing: every pass overwrites what a later pass reads.
static constexpr size_t kNoallocScratchBytes = temp_bytes_for(kNoallocMaxSlots);
static __device__ __align__(256) uint8_t scratch[kNoallocScratchBytes];

// Device address of `scratch` on the current device. In host code the array's
// name is only a stand-in for the symbol, so kernels get the address that
// hipGetSymbolAddress returns, looked up once per device with that device
// current.
static hipError_t noalloc_scratch(void **out) {
    static std::mutex mu;
    static void *ptr[kNoallocMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNoallocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        ptr[dev] = p;
    }
    *out = ptr[dev];
    return hipSuccess;
}

hipError_t launch_grid_build_noalloc(GridBuildLaunch g, hipStream_t s) {
    if (g.n_slots > kNoallocMaxSlots || temp_bytes_for(g.n_slots) > kNoallocScratchBytes)
        return hipErrorInvalidValue;  // the caller checks the size first and reports it
    void *p = nullptr;
    const hipError_t e = noalloc_scratch(&p);
    if (e != hipSuccess) return e;
    g.temp = p;
    g.temp_bytes = kNoallocScratchBytes;
    g.keys_in = nullptr;
    return launch_grid_build(g, s);
}

}  // namespace rtk
