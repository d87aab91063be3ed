// kernels.h — launch wrappers of the gfx950 kernels (host-callable).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "../../include/rt_types.h"
#include "../../include/rustracer_hip.h"

namespace rtk {

// Per-lane traversal stack capacity in 32-bit words (LDS rows + private spill).
constexpr uint32_t kTraversalStackWords = 96;

// The values of one frame that change from frame to frame, kept in device
// memory for replayed frame graphs: kernels given a FrameState read the camera,
// the seed frame and the accumulation weight from it instead of their
// arguments, so one captured graph serves every frame.
struct FrameState {
    rt_camera_uniform cam;
    uint32_t frame;  // ImageParams.frame (the seed, gather.wgsl:262)
    float alpha;     // accumulation weight of the frame (lib.rs:679-684)
    uint32_t pad[2];
};
static_assert(sizeof(FrameState) == 112, "FrameState layout");
hipError_t launch_set_frame_state(FrameState *dst, const FrameState &v, hipStream_t s);

// Traversal inputs shared by every tracing kernel (device pointers).
struct TraceInputs {
    const rt_bvh_node *nodes;
    const uint32_t *prims;
    const void *slot_tris;  // float4[3 * n_slots]
    const void *wnodes;     // float4[4 * n_internal]
    const void *w4nodes;    // float4[8 * n_records] (4-wide)
    uint32_t n_nodes;
    float root_min[4];      // w: root ref bits
    float root_max[4];
};

struct GatherLaunch {
    TraceInputs trace;
    const rt_triangle *tris;
    const rt_material *mats;
    const rt_light *lights;
    uint32_t n_lights;
    uint32_t n_tris;
    rt_camera_uniform cam;
    rt_image_params ip;
    rt_grid_params gp;
    const uint32_t *grid_meta;  // [start, count] per cell
    const rt_photon *sorted_photons;
    const void *gather_rec;     // 3 float4 per sorted photon: the radius gather's record
    void *frame;    // float4 per pixel
    void *moments;  // float4 per pixel
    void *gbuffer;  // float4 per pixel
    uint32_t max_depth;
    uint32_t mode;           // 0 reference nodes, 1 child-box reference order, 2 child-box nearest first, 3 4-wide
    uint32_t early_shadow;   // shadow rays stop at the first hit nearer than the light
    unsigned long long *counters;  // non-null: count traffic (4 counters)
    // path queues: when px_state is non-null each sample runs as a primary pass
    // plus one pass per depth over the queued paths (gather_wavefront_bytes: sizes)
    rt_tex_array tex[RT_TEX_ROLES];  // texture arrays (device texels)
    void *px_state;          // 5 * n_px float4
    void *queue[2];          // 3 * n_px float4 each; queue[b & 1]: the paths entering depth b
    uint32_t *path_count;    // 3 per depth b >= 1 at 3 (b - 1): work taken, diffuse paths queued, smooth paths queued
    // persistent form: when non-null (and px_state is null) the frame runs as
    // resident wavefronts whose lanes take pixels from 8 band counters here
    uint32_t *regen;
    uint32_t regen_blocks;   // cap on its workgroups (0 = all that are resident at once)
    uint32_t refill_min;     // a wavefront takes new pixels once this many of its lanes have none
    const FrameState *fstate;  // non-null: camera and seed frame come from here
    // persistent kernel: consecutive frames rendered by this one launch (0 or 1
    // = one). Frame f of the launch has seed frame ip.frame + f and writes
    // frame/moments/gbuffer + f * frame_stride float4 elements.
    uint32_t n_frames;
    uint32_t frame_stride;
};
// Device bytes of the path-queue state for n_px pixels.
size_t gather_wavefront_bytes(uint32_t n_px, uint32_t max_depth);

hipError_t launch_gather(const GatherLaunch &g, hipStream_t s);

// One view of a batch as the batch kernel reads it (128 bytes).
struct BatchViewRecord {
    rt_camera_uniform cam;
    uint32_t width, height, spp, frame;
    uint64_t pixel_offset;
    uint32_t tiles_x;
    uint32_t pad;
};
struct BatchLaunch {
    const BatchViewRecord *views;
    const uint32_t *tiles;     // (view, tile) pairs
    uint32_t n_tiles;
    uint32_t tile_w, tile_h;
    uint32_t tiles_per_fetch;
    uint32_t n_workgroups;     // persistent workgroups (one wavefront each)
    uint32_t *counter;
    uint32_t lanes;            // 1: lanes take pixel slots one at a time; 0: wavefronts claim tiles
};
// g.frame / g.moments / g.gbuffer point at the batch outputs (sum of view pixels).
hipError_t launch_gather_batch(const GatherLaunch &g, const BatchLaunch &b, hipStream_t s);
// Slot triangles of the watertight traversal (mode 4): (v0, prim bits), v1, v2
// of triangle prims[k] for slot k, copied from the uploaded triangles.
hipError_t launch_wt_slots(const rt_triangle *tris, const uint32_t *prims, uint32_t n_slots, void *out, hipStream_t s);
hipError_t launch_trace_rays(const TraceInputs &t, uint32_t mode, const rt_ray *rays, uint32_t n, rt_hit *out,
                             hipStream_t s);

struct PhotonLaunch {
    TraceInputs trace;
    const rt_triangle *tris;
    const rt_material *mats;
    const rt_light *lights;
    uint32_t n_lights;
    uint32_t n_tris;
    rt_grid_params gp;
    uint32_t n_photons;  // invocations (photon indices 0..n-1)
    uint32_t max_depth;  // slots per photon
    uint32_t mode;
    rt_photon *photons;
    uint32_t *keys = nullptr;  // n_photons * max_depth cell keys, pre-filled with num_cells; may be null
    rt_tex_array tex[RT_TEX_ROLES];
};
hipError_t launch_photon_trace(const PhotonLaunch &p, hipStream_t s);

// Grid build: keys, compaction of the valid slots, stable radix sort of (key, slot), grid_meta, gather.
struct GridBuildLaunch {
    const rt_photon *photons;
    uint32_t n_slots;      // slots processed
    rt_grid_params gp;
    uint32_t meta_cells;   // >= num_cells
    uint32_t *grid_meta;   // 2 per cell
    rt_photon *sorted;
    void *gather_rec;      // 3 float4 per sorted photon (position, power, incident)
    // temporaries (device), sized by grid_build_temp_bytes
    void *temp;
    size_t temp_bytes;
    uint32_t *n_valid_out; // device u32
    const uint32_t *keys_in = nullptr;  // cell key per slot from the photon trace; null: computed from photons
};
size_t grid_build_temp_bytes(uint32_t n_slots, uint32_t meta_cells);
hipError_t launch_grid_build(const GridBuildLaunch &g, hipStream_t s);
// The same build from uploaded photons (keys computed from the records) with
// the temporaries in the library's fixed per-device workspace instead of
// g.temp: no allocation. n_slots <= kNoallocMaxSlots; gather_rec may be null.
constexpr uint32_t kNoallocMaxSlots = RT_GRID_NOALLOC_MAX_SLOTS;
constexpr int kNoallocMaxDevices = 64;
hipError_t launch_grid_build_noalloc(GridBuildLaunch g, hipStream_t s);

struct DenoiseLaunch {
    const void *in;
    void *out;
    const void *gbuffer;
    const void *moments;
    rt_denoise_params p;
    const FrameState *fstate;  // non-null: alpha comes from here
};
hipError_t launch_denoise(const DenoiseLaunch &d, hipStream_t s);

// fstate non-null: alpha comes from fstate->alpha
// hist = hist * (1 - alpha) + moments.xy * alpha per pixel (the accumulation's
// weighting, applied to the frame's luminance moments); fstate non-null: alpha
// comes from fstate->alpha.
hipError_t launch_moment_history(const void *moments, void *hist, uint32_t n_px, float alpha,
                                 const FrameState *fstate, hipStream_t s);
hipError_t launch_composite_tonemap(const void *frame, void *accum, uint32_t *rgba8, uint32_t w, uint32_t h,
                                    float alpha, const rt_tonemap_uniform &tm, const void *gbuffer,
                                    const FrameState *fstate, hipStream_t s);
hipError_t launch_tonemap(const void *img, uint32_t *rgba8, uint32_t w, uint32_t h, const rt_tonemap_uniform &tm,
                          const void *gbuffer, hipStream_t s);
// The present pass: rgba8 output (tw x th, tight rows) resampled into a bgra8
// surface of sw x sh with pitch_words 32-bit words between rows.
hipError_t launch_present(const uint32_t *rgba8, uint32_t tw, uint32_t th, uint32_t *surface, uint32_t sw,
                          uint32_t sh, uint32_t pitch_words, hipStream_t s);
// First-hit albedo of the pixel-centre camera ray (rgb; w = 1 on a hit, 0 on a miss).
hipError_t launch_albedo(const TraceInputs &t, uint32_t mode, const rt_triangle *tris, const rt_material *mats,
                         const rt_camera_uniform &cam, uint32_t w, uint32_t h, void *out, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s);
hipError_t launch_accum64_add(const void *accum_f4, double *sum, size_t n_px, hipStream_t s);
hipError_t launch_hbm_copy(const void *src, void *dst, size_t bytes, hipStream_t s);
// relMSE of a float4 image against sum / n_frames (double4 per pixel); the
// result lands in scratch[relmse_scratch_doubles() - 1].
size_t relmse_scratch_doubles();
hipError_t launch_relmse(const void *img, const double *sum, uint64_t n_frames, float eps, size_t n_px,
                         double *scratch, hipStream_t s);

}  // namespace rtk
