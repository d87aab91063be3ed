/*
 * rustracer_hip.h — C ABI of the gfx950 (MI355X) backend for rustracer's
 * rendering path.
 *
 * The reference boundary is the Rust crate API of `rustracer-renderer`
 * (rustracer-renderer/src/lib.rs) plus the scene model of `rustracer-core`.
 * wgpu types cannot cross a C ABI, so the adapter becomes a HIP device ordinal,
 * `record_frame(encoder, view)` becomes rt_render_frame(), and read-backs
 * (examples/common/mod.rs) become copy-out calls. Every entry point returns an
 * int status (RT_OK = 0); on failure rt_last_error() returns a message for the
 * calling thread. Host pointers are borrowed for the duration of the call and
 * copied. A renderer handle owns its device memory and one HIP stream and is
 * not thread-safe (callers serialise, as `&mut self` does in the reference).
 * No HIP or torch types appear in any signature.
 */
#ifndef RUSTRACER_HIP_H
#define RUSTRACER_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "rt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RT_OK 0
#define RT_ERR_INVALID -1     /* bad argument / null pointer / size mismatch */
#define RT_ERR_UNSUPPORTED -2 /* feature outside this backend (e.g. textures) */
#define RT_ERR_HIP -3         /* HIP runtime error (message in rt_last_error) */
#define RT_ERR_NO_DEVICE -4   /* no usable gfx950 device */
#define RT_ERR_CAPACITY -5    /* a size exceeds a documented limit */

/* Message of the last failing call on this thread ("" if none). */
const char *rt_last_error(void);
/* Library version string. */
const char *rt_version(void);

/* ------------------------------------------------------------------------ */
/* Render configuration. Replaces RenderConfig, rustracer-renderer/src/lib.rs:21-65.
 * Fields up to `ambient` mirror the reference one to one (bool -> uint32 0/1).
 * The extension fields default to the reference's fixed behaviour.           */
typedef struct rt_config {
    uint32_t width;
    uint32_t height;
    uint32_t photon_count;
    uint32_t spp;
    float exposure;
    float gamma;
    float light_emission[3];
    float point_light_emission[3];
    uint32_t accumulate_frames; /* 0 = overwrite, N = running mean then EMA over N */
    uint32_t enable_photon_map;
    uint32_t enable_denoise;
    uint32_t photon_debug;
    uint32_t debug_view;
    float photon_scale;
    float photon_radius; /* overwritten by grid sizing at creation, lib.rs:431 */
    float ambient;
    /* --- extensions (no counterpart in RenderConfig) --- */
    uint32_t max_depth;     /* bounce limit of camera and photon paths; reference: 6 */
    float grid_cell_size;   /* 0: reference sizing (max_extent/24, <= 26 cells/axis);
                               > 0: this cell size and radius, cells sized to the scene */
    uint32_t denoise_passes;/* a-trous passes at steps 1,2,4..; reference: 2 */
    uint32_t flags;         /* RT_FLAG_* */
} rt_config;

/* Time each pass with HIP events on the renderer's stream (rt_get_pass_stats). */
#define RT_FLAG_PROFILE 1u
/* Trace shadow rays with the reference's closest-hit query instead of the
 * equivalent early-exit query (same result when no material is alpha-cutout). */
#define RT_FLAG_CLOSEST_SHADOW 2u
/* Traverse the reference's 48-byte BVH2 nodes instead of the child-box records
 * (both give the reference's hits bit for bit). */
#define RT_FLAG_REFERENCE_TRAVERSAL 4u
/* Child-box traversal nearest child first: fewer nodes visited; equals the
 * reference except where a hit's rounded t falls before its own box's entry. */
#define RT_FLAG_NEAREST_FIRST 8u
/* Use the binary child-box records instead of the default 4-wide records
 * (both exact; for measurement). */
#define RT_FLAG_BINARY_RECORDS 16u
/* Run the camera pass as path-queue passes (per sample a primary pass, then one
 * pass over the compacted live paths, sorted by the kind of their first
 * scattering: diffuse, or specular/dielectric) instead of the default
 * persistent kernel (same results). */
#define RT_FLAG_PATH_QUEUES 32u
/* Run the camera pass as one workgroup per 8x8 tile, each lane running one
 * pixel to the end, instead of the default resident wavefronts whose lanes take
 * a new pixel whenever theirs is done (same results). */
#define RT_FLAG_TILE_LAUNCH 64u
/* Replay each frame from hipGraphs captured once per output set: the gather
 * pass, and the a-trous + composite + tonemap passes. The camera, seed frame
 * and accumulation weight are written to device memory ahead of each replay;
 * a graph is captured again when anything else it was made from changes
 * (config, photon map, textures, buffers). Same results. */
#define RT_FLAG_GRAPH 128u
/* Keep a running per-pixel history of the frames' luminance moments (E[l],
 * E[l^2], weighted as the accumulation is) beside the accumulation. The learned
 * denoiser's variance features read it; it is kept whenever a denoiser is
 * attached, and this flag keeps it without one (for rt_unet_dataset_set). No
 * counterpart in the reference, whose denoiser sees only one frame's moments. */
#define RT_FLAG_MOMENT_HISTORY 256u
/* Test triangles with the watertight ray/triangle test (Woop, Benthin, Wald
 * 2013) instead of the reference's Moller-Trumbore, on the 4-wide records
 * (traversal mode 4): a ray through an edge or vertex shared by two triangles
 * hits one of them. Not bit-exact with the reference: a hit's t differs in the
 * last bits and edge rays may pick the other triangle of the edge (counted by
 * the tests). No counterpart in the reference. */
#define RT_FLAG_WATERTIGHT 512u

/* Fill `cfg` with RenderConfig::default() (lib.rs:43-64) plus extension defaults. */
int rt_config_default(rt_config *cfg);

/* ------------------------------------------------------------------------ */
/* Scene model (rustracer-core). Replaces Scene, rustracer-core/src/scene.rs:158-291. */
typedef struct rt_scene rt_scene;

/* Light description (rustracer-core Light enum, scene.rs:101-126). */
typedef struct rt_light_desc {
    uint32_t kind;          /* RT_LIGHT_* */
    float vec_a[3];         /* point: position; directional: direction; environment: zenith colour */
    float vec_b[3];         /* environment: horizon colour */
    float color[3];         /* point/directional/area colour */
    float intensity;
    uint32_t triangle_index;/* area: index into the flattened triangle list */
} rt_light_desc;

/* Host camera (rustracer-core Camera, camera.rs:10-32). */
typedef struct rt_camera {
    float position[3];
    float look_at[3];
    float up[3];
    float fov_degrees;
    float near_plane;
    float aperture;
    float focus_dist;
} rt_camera;

int rt_scene_new(rt_scene **out);
void rt_scene_free(rt_scene *scene);
/* Append a material record; returns its index in *index (Material::to_gpu, material.rs:143-161). */
int rt_scene_add_material(rt_scene *scene, const rt_material *material, uint32_t *index);
/* Append an indexed mesh (Mesh::new, scene.rs:24-38). normals/uvs may be NULL
 * (flatten_triangles substitutes +Y / zero, scene.rs:233-250). */
int rt_scene_add_mesh(rt_scene *scene, const float *positions, uint32_t n_positions,
                      const float *normals, uint32_t n_normals, const float *uvs,
                      uint32_t n_uvs, const uint32_t *indices, uint32_t n_indices,
                      uint32_t material_id);
int rt_scene_add_light(rt_scene *scene, const rt_light_desc *light);
int rt_scene_clear_lights(rt_scene *scene);
/* Built-in procedural scenes, seeded:
 *   "cornell_box"   Scene::cornell_box (scene.rs:296-381), 60 triangles
 *   "dark_room"     Scene::dark_room (scene.rs:386-460)
 *   "smoke"         the 3-triangle scene of gpu_layout.rs:66-99
 *   "cornell36"     36-triangle unit Cornell room, radiance-15 ceiling quad
 *   "photon_plain"  displaced icosphere (~46k) + glass sphere (8k) + floor + area and point light
 *   "multiview"     ~109k-triangle tori / bevelled boxes, Lambert/GGX/glass mix
 *   "fullframe"     ~1M triangles: value-noise terrain (600k) + 200 meshes of 2k
 *   "training"      one seeded procedural training scene (instanced primitives)   */
int rt_scene_builtin(const char *name, uint64_t seed, rt_scene **out);
/* Camera suited to a built-in scene (index selects one of several for "multiview"). */
int rt_scene_builtin_camera(const char *name, uint64_t seed, uint32_t index, rt_camera *out);

int rt_scene_counts(const rt_scene *scene, uint32_t *n_triangles, uint32_t *n_materials,
                    uint32_t *n_lights, uint32_t *n_meshes);
/* Scene::bounds (scene.rs:198-210); returns RT_ERR_INVALID when there is no geometry. */
int rt_scene_bounds(const rt_scene *scene, float bmin[3], float bmax[3]);
/* Scene::flatten_triangles (scene.rs:213-273). */
int rt_scene_flatten_triangles(const rt_scene *scene, rt_triangle *out, uint32_t capacity);
int rt_scene_flatten_materials(const rt_scene *scene, rt_material *out, uint32_t capacity);
/* Scene::tex_albedo / tex_mr / tex_emissive / tex_normal (rustracer-core
 * scene.rs, texture.rs): append a width x height RGBA float texture to a role
 * (RT_TEX_ROLE_*); *layer receives its layer index (the material's tex_* value). */
int rt_scene_add_texture(rt_scene *scene, uint32_t role, uint32_t width, uint32_t height, const float *rgba,
                         uint32_t *layer);
/* The role's rgba8 array as the renderer uploads it (lib.rs:1286-1358); with
 * texels = NULL only the sizes are returned. */
int rt_scene_texture_array(const rt_scene *scene, uint32_t role, uint8_t *texels, size_t capacity, uint32_t *width,
                           uint32_t *height, uint32_t *layers);
/* light_to_gpu over the scene's lights (lib.rs:1231-1282). */
int rt_scene_lights_gpu(const rt_scene *scene, rt_light *out, uint32_t capacity);
/* The scene's light descriptions; *n = their count (out = NULL: count only). */
int rt_scene_light_descs(const rt_scene *scene, rt_light_desc *out, uint32_t capacity, uint32_t *n);
/* Parse a scene.toml descriptor and apply it to the scene, as
 * parse_descriptor + apply_descriptor do (rustracer-core/src/config.rs:62-174):
 * [environment] replaces every environment light, [[lights]] of type point,
 * directional or area are appended (other types skipped), [[materials]]
 * overrides the material at `index`, or every material without one (indices
 * out of range skipped). A malformed file or a field of the wrong type returns
 * RT_ERR_INVALID and leaves the scene unchanged. */
int rt_scene_apply_descriptor(rt_scene *scene, const char *toml_text);
/* The PNG tonemap of render_headless (examples/common/mod.rs:37-47) over n_px
 * float4 pixels: rgb -> u8 of (v / (1 + v))^(1/2.2) * 255, truncated; alpha 255. */
int rt_tonemap_cpu(const float *accum_rgba, size_t n_px, uint8_t *rgba8);
/* The glTF loader's sRGB decode (loader.rs:28-34) in place, in f32 with the C
 * library's powf: c <= 0.04045 ? c / 12.92 : ((c + 0.055) / 1.055)^2.4. */
int rt_srgb_to_linear(float *values, size_t n);
int rt_light_to_gpu(const rt_light_desc *light, rt_light *out);
/* Scene::build_bvh (scene.rs:276-286) -> BVHBuilder::build (bvh.rs:135-151):
 * binned SAH (12 bins, <= 4 prims per leaf), root at node 0. Call with NULL
 * outputs to query the sizes. */
int rt_scene_build_bvh(const rt_scene *scene, rt_bvh_node *nodes, uint32_t node_capacity,
                       uint32_t *n_nodes, uint32_t *prims, uint32_t prim_capacity,
                       uint32_t *n_prims);
/* BVHBuilder over explicit centroids/boxes (bvh.rs:127-151), for the builder tests. */
int rt_bvh_build_boxes(const float *centroids, const float *box_min, const float *box_max,
                       uint32_t n, rt_bvh_node *nodes, uint32_t node_capacity,
                       uint32_t *n_nodes, uint32_t *prims);

/* Camera (camera.rs). */
int rt_camera_default(rt_camera *cam);                                              /* :42-52  */
int rt_camera_to_uniform(const rt_camera *cam, uint32_t w, uint32_t h,
                         rt_camera_uniform *out);                                   /* :56-79  */
int rt_camera_frame_scene(const rt_scene *scene, rt_camera *out);                   /* :83-99  */
int rt_camera_orbit(rt_camera *cam, float delta_yaw, float delta_pitch);            /* :102-115 */
int rt_camera_zoom(rt_camera *cam, float delta);                                    /* :118-123 */
int rt_camera_pan(rt_camera *cam, float dx, float dy);                              /* :126-133 */

/* Photon-grid sizing of Renderer::new (lib.rs:416-457), or the free-cell-size
 * extension when cfg->grid_cell_size > 0. Writes the grid and the radius the
 * renderer will use. */
int rt_grid_params_for(const rt_config *cfg, const float bmin[3], const float bmax[3],
                       int has_bounds, rt_grid_params *out, float *photon_radius);

/* ------------------------------------------------------------------------ */
/* Renderer (rustracer-renderer Renderer, lib.rs:189-1041). */
typedef struct rt_renderer rt_renderer;

/* Renderer::new (lib.rs:227-663): flatten, build the BVH, upload, size the grid. */
int rt_renderer_create(const rt_config *cfg, const rt_scene *scene, int device,
                       rt_renderer **out);
/* Renderer::new over already-flattened GPU records (the upload contract of
 * lib.rs:234-271). lights may be empty (a zeroed record is bound, buffer.rs:19-34).
 * bounds: the scene bounds used for grid sizing (NULL = no geometry). */
int rt_renderer_create_raw(const rt_config *cfg, const rt_triangle *tris, uint32_t n_tris,
                           const rt_bvh_node *nodes, uint32_t n_nodes, const uint32_t *prims,
                           uint32_t n_prims, const rt_material *mats, uint32_t n_mats,
                           const rt_light *lights, uint32_t n_lights, const float *bounds_min,
                           const float *bounds_max, int device, rt_renderer **out);
void rt_renderer_destroy(rt_renderer *r);
/* The texture array of one role (RT_TEX_ROLE_*), rgba8 texels layers x height x
 * width, as create_role_array uploads it (lib.rs:1286-1358); for renderers made
 * with rt_renderer_create_raw (rt_renderer_create takes the scene's textures). */
int rt_renderer_set_texture_array(rt_renderer *r, uint32_t role, uint32_t width, uint32_t height, uint32_t layers,
                                  const uint8_t *texels);

/* record_frame + queue.submit (lib.rs:666-851): gather, a-trous, composite,
 * tonemap. Asynchronous; rt_synchronize waits (device.poll(Wait)). */
int rt_render_frame(rt_renderer *r);
/* n_frames record_frame calls in a row (the frame loop of render_headless.rs:
 * 132-136): the same frames and results, with up to eight frames rendered by
 * one camera-pass launch. */
int rt_render_frames(rt_renderer *r, uint32_t n_frames);
int rt_synchronize(rt_renderer *r);
int rt_update_camera(rt_renderer *r, const rt_camera *cam);                 /* lib.rs:853-857 */
int rt_update_camera_uniform(rt_renderer *r, const rt_camera_uniform *u);
int rt_update_lights(rt_renderer *r, const rt_light *lights, uint32_t n);   /* lib.rs:888-891 */
int rt_reset_accumulation(rt_renderer *r);                                  /* lib.rs:893-902 */
int rt_reset_accumulation_soft(rt_renderer *r);                             /* lib.rs:907-909 */
int rt_rebuild_photons(rt_renderer *r);                                     /* lib.rs:913-1037 */
uint32_t rt_sample_count(const rt_renderer *r);                             /* lib.rs:1039-1041 */
/* Set the counter so that the next frame uses seed frame index `value + 1`. */
int rt_set_sample_count(rt_renderer *r, uint32_t value);
/* The mutable `pub config` (main.rs:505,545): width/height/photon_count/max_depth
 * are fixed at creation and must not change. */
int rt_get_config(const rt_renderer *r, rt_config *out);
int rt_set_config(rt_renderer *r, const rt_config *cfg);
int rt_get_grid_params(const rt_renderer *r, rt_grid_params *out);

/* Read-backs (examples/common/mod.rs:6-30) and parity hooks. Sizes in elements. */
int rt_read_accumulation(rt_renderer *r, float *rgba, size_t n_floats);
int rt_read_frame(rt_renderer *r, float *rgba, size_t n_floats);
int rt_read_moments(rt_renderer *r, float *rgba, size_t n_floats);
int rt_read_gbuffer(rt_renderer *r, float *rgba, size_t n_floats);
int rt_read_output_rgba8(rt_renderer *r, uint8_t *rgba, size_t n_bytes);
int rt_write_accumulation(rt_renderer *r, const float *rgba, size_t n_floats);

/* The present pass of Renderer::record_frame (lib.rs:834-850, fullscreen.wgsl:
 * the fullscreen-triangle render pass into the surface_view argument of
 * record_frame, lib.rs:667-670): the tonemapped rgba8 output, sampled with
 * ClampToEdge bilinear filtering (lib.rs:555-562), written as Bgra8Unorm pixels
 * (the app's surface format, rustracer-app/src/main.rs:320-324) into a surface
 * of surface_w x surface_h that need not have the render size. d_surface is
 * device memory with row_pitch bytes between rows (a multiple of 4, at least
 * 4*surface_w); the pass is queued on the renderer's stream after the frames
 * already issued and returns without waiting. Sides are 1..32768. */
int rt_present(rt_renderer *r, void *d_surface, uint32_t surface_w, uint32_t surface_h, size_t row_pitch);
/* rt_present into a surface buffer the renderer owns (allocated on the first
 * call of a larger size), read back to the host with tight rows. */
int rt_read_present(rt_renderer *r, uint32_t surface_w, uint32_t surface_h, uint8_t *bgra, size_t n_bytes);
int rt_read_photons(rt_renderer *r, rt_photon *out, size_t capacity, uint32_t *n_slots);
int rt_read_sorted_photons(rt_renderer *r, rt_photon *out, size_t capacity, uint32_t *n_stored);
int rt_read_grid_meta(rt_renderer *r, uint32_t *meta, size_t n_uint32, uint32_t *n_cells);

/* Closest-hit queries over the uploaded BVH for a fixed ray set (the trace_bvh
 * of gather.wgsl:66-88 as an entry point). Host arrays. */
int rt_trace_closest(rt_renderer *r, const rt_ray *rays, uint32_t n, rt_hit *out);
/* Same, with ray and hit arrays already in device memory, on the renderer's stream. */
int rt_trace_closest_device(rt_renderer *r, const rt_ray *d_rays, uint32_t n, rt_hit *d_out);
/* Which traversal the renderer runs (0 reference BVH2 nodes, 1 binary
 * child-box records, 2 binary nearest-first, 3 4-wide records), the BVH depth
 * and the number of 4-wide records (0 when that layout is unavailable). No
 * reference counterpart: it reports this backend's derived layouts. */
int rt_traversal_info(const rt_renderer *r, uint32_t *mode, uint32_t *bvh_depth, uint32_t *n_records4);
/* Worst-case per-lane traversal stack words of the traversal in use, from the
 * tree's shape (no counterpart in the reference). */
int rt_traversal_stack_words(const rt_renderer *r, uint32_t *words);
/* The light gizmo of the tonemap pass (pub light_marker / light_glow,
 * lib.rs:201-203, drawn by tonemap.wgsl:32-47): a warm ball of world radius
 * `radius` at `position`, hidden behind nearer geometry. position = NULL hides
 * it (light_marker = None). Projected every frame with project_marker
 * (lib.rs:865-885) from the Camera given to rt_update_camera; a camera given
 * only as a uniform (rt_update_camera_uniform) hides it. */
int rt_set_light_marker(rt_renderer *r, const float *position, float radius);
int rt_set_light_glow(rt_renderer *r, float glow);
/* Renderer::project_marker for a camera and image size: out = (centre x, y in
 * pixels, radius in pixels, camera distance or 0 when hidden). */
int rt_project_marker(const rt_camera *cam, uint32_t width, uint32_t height, const float position[3], float radius,
                      float out[4]);

/* Number of frame graphs captured so far (RT_FLAG_GRAPH): two per output set
 * at first use, then one more each time a captured pass's inputs change. No
 * counterpart in the reference. */
int rt_frame_graph_captures(const rt_renderer *r, uint32_t *captures);
/* Build the photon grid from a given photon array (count, scan, stable
 * scatter: photon_count.wgsl, lib.rs:987-1007, photon_scatter.wgsl) into the
 * renderer's grid, replacing traced photons. n_slots <= photon_count*max_depth. */
int rt_build_photon_grid_from(rt_renderer *r, const rt_photon *photons, uint32_t n_slots);
/* Allocation-free variant of rt_build_photon_grid_from. It has no counterpart
 * in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It builds the grid of n_slots photon records already in device
 * memory (d_photons) into caller-owned device outputs, on `stream` (a
 * hipStream_t of the current device; NULL = the null stream), asynchronously:
 *   d_grid_meta  [start, count] per cell, meta_cells >= gp->num_cells entries
 *   d_sorted     the stored photons in cell order, stable (n_slots records fit)
 *   d_gather_rec 3 float4 per stored photon (position, power, incident), or NULL
 *   d_n_sorted   one u32: the number of stored photons
 * The temporaries live in a fixed library-owned device array for
 * RT_GRID_NOALLOC_MAX_SLOTS slots (about 17 MB), one per GPU, held until the
 * process ends; larger inputs fail with RT_ERR_CAPACITY. Calls on one GPU share
 * that array, so calls that overlap on one GPU race. The first call on a device
 * looks up the array's address (make it before capturing a graph). */
#define RT_GRID_NOALLOC_MAX_SLOTS (1u << 20)
int rt_photon_grid_build_noalloc(const rt_photon *d_photons, uint32_t n_slots, const rt_grid_params *gp,
                                 uint32_t *d_grid_meta, uint32_t meta_cells, rt_photon *d_sorted,
                                 float *d_gather_rec, uint32_t *d_n_sorted, void *stream);

/* Per-pass timing with HIP events (cfg.flags & RT_FLAG_PROFILE). */
typedef struct rt_pass_stats {
    double ms[12];       /* indexed by RT_PASS_*; unused entries stay 0 */
    uint64_t launches[12];
} rt_pass_stats;
#define RT_PASS_GATHER 0
#define RT_PASS_DENOISE 1
#define RT_PASS_COMPOSITE_TONEMAP 2
#define RT_PASS_PHOTON_TRACE 3
#define RT_PASS_GRID_BUILD 4
#define RT_PASS_TRACE_RAYS 5
#define RT_PASS_ACCUM64 6
int rt_get_pass_stats(rt_renderer *r, rt_pass_stats *out);
int rt_reset_pass_stats(rt_renderer *r);

/* Traversal counters of one extra, untimed gather launch of the current frame
 * state (rays, BVH nodes loaded, triangle records loaded, photon records read),
 * for the algorithmic-bytes figure of the roofline. */
#define RT_COUNT_PHASES 5
typedef struct rt_trace_counters {
    uint64_t ray_queries;
    uint64_t nodes_visited;
    uint64_t tris_tested;
    uint64_t photon_reads;
    /* loop steps per phase, summed over lanes and over wavefronts:
     * lane_steps / wave_steps = mean active lanes of 64 in that phase.
     * Phases: 0 traversal steps of closest-hit queries, 1 one bounce body,
     * 2 one batch of photon candidates, 3 traversal steps of shadow queries,
     * 4 one iteration of the persistent camera pass's loop (live lanes) */
    uint64_t lane_steps[RT_COUNT_PHASES];
    uint64_t wave_steps[RT_COUNT_PHASES];
    /* shader-clock cycles wavefronts spent in each phase's code regions, summed
     * over wavefronts (phase 1: whole bounces, their traversals and photon
     * search included; phase 4: whole iterations of the persistent loop) */
    uint64_t wave_cycles[RT_COUNT_PHASES];
} rt_trace_counters;
int rt_count_gather_traffic(rt_renderer *r, rt_trace_counters *out);

/* ------------------------------------------------------------------------ */
/* Many views in one persistent launch. No counterpart in the reference, which
 * renders one camera per dispatch: each view is the gather of gather.wgsl with
 * its own camera, resolution, spp and seed frame (every other parameter comes
 * from the renderer's config), and its pixels are bit-identical to a
 * single-view render with those values. Persistent workgroups pull work from a
 * global queue of tiles (most expensive first): whole tiles per wavefront, or
 * pixel slots per lane, each lane taking the next pixel when its own is done. */
typedef struct rt_view {
    rt_camera_uniform camera; /* Camera::to_uniform at this view's resolution */
    uint32_t width;
    uint32_t height;
    uint32_t spp;
    uint32_t frame; /* ImageParams.frame of the seed, gather.wgsl:262 */
} rt_view;

typedef struct rt_batch_tuning {
    uint32_t tile_w, tile_h;  /* pixels per tile; tile_w * tile_h a multiple of 64 */
    uint32_t waves_per_cu;    /* persistent workgroups (one wavefront each) per CU, 1..20 */
    uint32_t tiles_per_fetch; /* tiles claimed per queue access (lanes = 0) */
    uint32_t lanes;           /* 1: lanes take pixel slots as they finish; 0: wavefronts claim whole tiles */
} rt_batch_tuning;

int rt_batch_set_views(rt_renderer *r, const rt_view *views, uint32_t n);
/* One launch over every view; NULL uses the current tuning (default 8x8 tiles, 16 waves/CU,
 * lanes taking pixel slots). */
int rt_render_batch(rt_renderer *r, const rt_batch_tuning *tuning);
/* Sweep tile shape, waves per CU and queue granularity (pixel slots per lane, or
 * 1 or 4 tiles per wavefront) on the current views,
 * keep the fastest (mean of `repeats` timed launches after one warm-up). */
int rt_batch_autotune(rt_renderer *r, uint32_t repeats, rt_batch_tuning *best, double *best_ms);
int rt_batch_get_tuning(const rt_renderer *r, rt_batch_tuning *out);
int rt_batch_totals(const rt_renderer *r, uint64_t *pixels, uint64_t *samples);
/* frame / moments / G-buffer of one view (each width*height*4 floats; NULL skips). */
int rt_read_batch_view(rt_renderer *r, uint32_t view, float *frame, float *moments, float *gbuffer,
                       size_t n_floats);
#define RT_PASS_BATCH 7
#define RT_PASS_NETWORK 8   /* learned denoiser: albedo pass, network, tonemap of its output */
#define RT_PASS_PRESENT 9   /* rt_present / rt_read_present */

/* Measured HBM bandwidth: a plain device-to-device copy kernel of `bytes`,
 * repeated `iters` times; *gbps counts bytes read + written. No counterpart in
 * the reference (measurement support). */
int rt_hbm_copy_probe(int device, size_t bytes, uint32_t iters, double *gbps, double *ms_per_copy);

/* Convergence support: FP64 running sum of the displayed frames and its mean. */
int rt_accum64_reset(rt_renderer *r);
int rt_accum64_add_accumulation(rt_renderer *r);
int rt_accum64_read_mean(rt_renderer *r, double *rgba, size_t n_doubles, uint64_t *n_frames);
/* relMSE = mean over pixels and rgb of (x - y)^2 / (y^2 + eps) of a float4
 * image x (device pointer, or the accumulation buffer when null) against the
 * FP64 mean y = sum / frames. No counterpart in the reference (the convergence
 * measurement). */
int rt_accum64_relmse(rt_renderer *r, const void *d_image, float eps, double *relmse);
/* The same for the displayed image: the learned denoiser's output when one is
 * attached, else the accumulation. */
int rt_accum64_relmse_display(rt_renderer *r, float eps, double *relmse);

/* ───────────── learned denoiser (no counterpart in the reference) ─────────────
 * A five-level convolutional U-Net (32-128 channels) on the displayed image:
 * 12 input channels, log1p(accumulated radiance), first-hit albedo
 * (pixel-centre ray), G-buffer normal and depth, and two variance channels from
 * the accumulated luminance moments; output a log-radiance residual. BF16 MFMA
 * convolutions with FP32 accumulation; trained on the GPU with Adam and an L1
 * loss on log radiance, on pairs the renderer itself produces (few-frame
 * accumulation -> many-sample target). Any image size (padded internally to
 * multiples of 16). Training is deterministic: fixed-order gradient sums. */
typedef struct rt_unet rt_unet;
int rt_unet_create(int device, uint64_t seed, rt_unet **out);
int rt_unet_destroy(rt_unet *u);
int rt_unet_param_count(const rt_unet *u, size_t *n);
int rt_unet_get_params(rt_unet *u, float *out, size_t n);
/* New parameters also restart the optimizer (Adam's moments and step count). */
int rt_unet_set_params(rt_unet *u, const float *in, size_t n);
/* Training set of `capacity` images of width x height (multiples of 32). */
int rt_unet_dataset_reset(rt_unet *u, uint32_t width, uint32_t height, uint32_t capacity);
/* Slot `slot` takes the renderer's current state: which = 0 the network input
 * (accumulation, G-buffer, albedo, moment history: render with
 * RT_FLAG_MOMENT_HISTORY), which = 1 the target (the accumulation). The
 * renderer's size must equal the dataset's. */
int rt_unet_dataset_set(rt_unet *u, uint32_t slot, int which, rt_renderer *r);
/* `steps` Adam steps on random crops (crop x crop, a multiple of 32) of the
 * slots that hold both; per-step L1 loss on log radiance into `losses` (nullable). */
int rt_unet_train(rt_unet *u, uint32_t steps, uint32_t batch, uint32_t crop, float lr, uint64_t seed,
                  float *losses);
/* One forward pass (and, with a target, the backward pass) on host tensors
 * (h, w multiples of 16): features n*h*w*16 bf16 bits, noisy and target n*h*w*3
 * log values; pred out n*h*w*3; grads out (param_count floats, same order as
 * the parameters); loss out (mean |pred - target|). */
int rt_unet_eval(rt_unet *u, uint32_t n, uint32_t h, uint32_t w, const uint16_t *features, const float *noisy,
                 const float *target, float *pred, float *grads, double *loss);
/* The convolution variant (input-channel slice width, output-channel blocks of
 * 32, weights staged once) picked for each layer shape so far: the first use of
 * a shape times every compiled variant that fits and keeps the fastest (all of
 * them give identical results). One text line per shape, "cin cout src N H W
 * mask bias -> kc nco once rpw", then one line per inference level whose pool
 * or upsample was timed folded against separate, "fold which N H W -> 1|0"
 * (which: 1 the level-1 pool, -1 - l decoder level l); *needed = bytes
 * including the terminator. */
int rt_unet_conv_choices(const rt_unet *u, char *buf, size_t cap, size_t *needed);
/* How a convolution variant is picked: 1 (default) the autotuned fastest, 0 the
 * first compiled for the shape, k >= 2 the (k-2)-th that fits (the last if
 * fewer): a test hook that runs every variant. */
int rt_unet_set_variant_mode(rt_unet *u, int mode);
/* Attach (or detach with NULL) a denoiser: each rt_render_frame then runs it on
 * the accumulation and tonemaps its output into the rgba8 image. */
int rt_renderer_set_denoiser(rt_renderer *r, rt_unet *u);
int rt_read_display(rt_renderer *r, float *rgba, size_t n_floats);
int rt_read_albedo(rt_renderer *r, float *rgba, size_t n_floats);

#ifdef __cplusplus
}
#endif

#endif /* RUSTRACER_HIP_H */
