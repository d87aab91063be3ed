// renderer.cpp — the Renderer of rustracer-renderer/src/lib.rs on HIP: device
// buffers, scene upload, per-frame pass sequence, photon-map rebuild,
// read-backs, parity hooks and per-pass timing. Exposed through include/rustracer_hip.h.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_util.h"
#include "kernels.h"
#include "scene.h"
#include "unet.h"

namespace {

using rt::ApiError;

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw ApiError(RT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T *dev_alloc(size_t count, const char *what) {
    void *p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hip_check(hipMalloc(&p, bytes), what);
    return static_cast<T *>(p);
}

constexpr uint32_t kTraversalStackWords = rtk::kTraversalStackWords;
// counting launch: queries, child boxes, triangles, photon candidates, then
// lane steps and wave steps per phase
constexpr int kCounterWords = 4 + 3 * RT_COUNT_PHASES;

struct PendingTiming {
    int pass;
    hipEvent_t start, stop;
};

}  // namespace

// What a captured post-pass graph (a-trous, composite, tonemap) was made from:
// a replay is valid while every field is unchanged.
struct PostKey {
    const void *frame, *ping, *moments, *gbuf, *accum, *rgba8;
    uint32_t width, height, denoise, passes;
    rt_tonemap_uniform tm;
};

struct rt_renderer {
    rt_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;

    uint32_t n_tris = 0, n_nodes = 0, n_prims = 0, n_mats = 0, n_lights = 0, n_internal = 0;
    rt_triangle *d_tris = nullptr;
    rt_bvh_node *d_nodes = nullptr;
    uint32_t *d_prims = nullptr;
    rt_material *d_mats = nullptr;
    rt_light *d_lights = nullptr;
    float *d_slot_tris = nullptr;     // float4[3 * n_prims]: v0, v1 - v0, v2 - v0
    float *d_slot_tris_wt = nullptr;  // float4[3 * n_prims]: v0, v1, v2 (watertight traversal)
    float *d_wnodes = nullptr;     // float4[4 * n_internal]
    float *d_w4nodes = nullptr;    // float4[8 * n_rec4]
    uint32_t n_rec4 = 0;
    bool wide4_ok = false;
    uint32_t stack_words4 = 0;  // worst-case traversal stack words of the 4-wide records
    float root_min[4] = {0, 0, 0, 0}, root_max[4] = {0, 0, 0, 0};
    bool wide_ok = false;
    uint32_t bvh_depth = 0;

    float *d_accum = nullptr, *d_frame = nullptr, *d_ping = nullptr, *d_moments = nullptr, *d_gbuf = nullptr;
    uint32_t *d_rgba8 = nullptr;
    // Frame pipelining: consecutive frames rotate through kSets output sets
    // (frame, a-trous ping-pong, moments, G-buffer) and alternate between
    // kStreams gather streams, so the camera pass of frame k+1 can start while
    // the last long paths of earlier frames are still running. The main stream
    // waits for each frame's gather before its a-trous and composite passes
    // (which stay in frame order), and a gather waits until the composite
    // kSets frames back has released its output set. d_frame etc. name the
    // latest frame's set. Each kind of buffer is one allocation of kSets
    // consecutive sets, so kChunk consecutive sets can be written by one
    // launch (rt_render_frames renders kChunk frames per camera pass launch).
    // measured on the 1080p workload (Msamples/s, rt_render_frames): chunks of
    // 2 frames 1385, 4 frames 1445, 8 frames 1485 (one frame per launch: 1165)
    static constexpr int kSets = 16;
    static constexpr int kStreams = 2;  // measured: three gathers in flight ran 10 % slower than two
    static constexpr int kChunk = 8;
    float *set_frame[kSets] = {}, *set_ping[kSets] = {}, *set_mom[kSets] = {}, *set_gb[kSets] = {};
    float *blk_frame = nullptr, *blk_ping = nullptr, *blk_mom = nullptr, *blk_gb = nullptr;
    hipStream_t gstream[kStreams] = {};
    hipEvent_t ev_gdone[kSets] = {}, ev_free[kSets] = {};
    bool free_recorded[kSets] = {};
    uint32_t next_set = 0;
    uint32_t chunk_stream = 0;

    rt_grid_params gp{};
    uint32_t meta_cells = 0;
    uint32_t photon_slots = 0;
    rt_photon *d_photons = nullptr, *d_sorted = nullptr;
    uint32_t *d_keys = nullptr;     // grid cell key per photon slot, written by the photon trace
    uint32_t *d_surface = nullptr;  // bgra8 surface of rt_read_present (allocated at the first use of a larger size)
    size_t surface_words = 0;
    float *d_gather_rec = nullptr;  // 3 float4 per sorted photon (radius-gather record)
    uint32_t *d_meta = nullptr;
    void *d_temp = nullptr;
    size_t temp_bytes = 0;
    uint32_t *d_nvalid = nullptr;
    uint32_t n_sorted = 0;
    bool photons_built = false;

    rt_camera_uniform cam{};
    // light gizmo (lib.rs:201-203, 704-707): drawn by the tonemap pass when a
    // marker is set and the camera came as a Camera (project_marker needs it)
    rt_camera host_cam{};
    bool host_cam_ok = false;
    bool marker_on = false;
    float marker_pos[3] = {0.f, 0.f, 0.f};
    float marker_radius = 0.f;
    float light_glow = 1.0f;
    uint32_t sample_counter = 0;
    uint32_t frames_since_reset = 0;

    unsigned long long *d_counters = nullptr;
    // pixel counters of the persistent camera pass: one block of 8 per output
    // set (gathers of two frames run at once) and one for the counting pass
    uint32_t *d_regen = nullptr;
    double *d_acc64 = nullptr;
    double *d_relmse = nullptr;  // reduction scratch of rt_accum64_relmse
    void *d_wave = nullptr;  // path-queue state of the camera pass
    size_t wave_cap = 0;
    // material texture arrays (lib.rs:1286-1358), one per role
    uint8_t *d_tex[RT_TEX_ROLES] = {};
    rt_tex_array tex[RT_TEX_ROLES] = {};
    bool has_cutout = false;  // an alpha-cutout material with an albedo texture
    // learned denoiser: attached network, its output image, the albedo guide
    rt_unet *unet = nullptr;
    float *d_display = nullptr;
    float *d_albedo = nullptr;
    bool albedo_ok = false;
    // running luminance moments (float2 per pixel), kept while a denoiser is
    // attached or RT_FLAG_MOMENT_HISTORY is set; weight of the latest post pass
    float *d_mom_hist = nullptr;
    float post_alpha = 1.0f;
    uint64_t acc64_frames = 0;

    // many-view batch
    std::vector<rt_view> views;
    std::vector<uint64_t> view_offsets;
    uint64_t batch_px = 0, batch_samples = 0;
    size_t batch_cap_px = 0;
    float *d_bframe = nullptr, *d_bmom = nullptr, *d_bgb = nullptr;
    rtk::BatchViewRecord *d_views = nullptr;
    size_t views_cap = 0;
    uint32_t *d_tiles = nullptr;
    size_t tiles_cap = 0;
    uint32_t n_tiles = 0, tiles_w = 0, tiles_h = 0;
    uint32_t *d_counter = nullptr;
    rt_batch_tuning tuning{8u, 8u, 16u, 1u, 1u};
    int n_cu = 0;

    std::vector<PendingTiming> pending;
    rt_pass_stats stats{};

    // Frame graphs (RT_FLAG_GRAPH): per output set, the gather pass and the post
    // passes captured once into hipGraphs and replayed every frame. The values
    // that change per frame (camera, seed frame, accumulation weight) are stored
    // into d_fstate[set] by a one-lane kernel ahead of each replay; a graph is
    // captured again only when what it was made from changes (gkey / pkey).
    rtk::FrameState *d_fstate = nullptr;
    hipStream_t cstream = nullptr;  // capture stream (nothing runs on it)
    hipGraphExec_t gexec[kSets] = {}, pexec[kSets] = {};
    rtk::GatherLaunch gkey[kSets] = {};
    PostKey pkey[kSets] = {};
    uint32_t graph_captures = 0;

    size_t n_px() const { return (size_t)cfg.width * cfg.height; }

    ~rt_renderer() {
        if (stream) hipStreamSynchronize(stream);
        for (int k = 0; k < kStreams; k++)
            if (gstream[k]) hipStreamSynchronize(gstream[k]);
        for (int k = 0; k < kSets; k++) {
            if (gexec[k]) hipGraphExecDestroy(gexec[k]);
            if (pexec[k]) hipGraphExecDestroy(pexec[k]);
        }
        if (cstream) hipStreamDestroy(cstream);
        for (int k = 0; k < kSets; k++) {
            if (ev_gdone[k]) hipEventDestroy(ev_gdone[k]);
            if (ev_free[k]) hipEventDestroy(ev_free[k]);
        }
        for (int k = 0; k < kStreams; k++)
            if (gstream[k]) hipStreamDestroy(gstream[k]);
        // d_frame etc. point into the set blocks, which are freed here
        for (float *b : {blk_frame, blk_ping, blk_mom, blk_gb})
            if (b) hipFree(b);
        d_frame = d_ping = d_moments = d_gbuf = nullptr;
        for (auto &p : pending) {
            hipEventDestroy(p.start);
            hipEventDestroy(p.stop);
        }
        void *bufs[] = {d_tris,  d_nodes, d_prims,   d_mats,    d_lights, d_slot_tris, d_slot_tris_wt, d_wnodes, d_accum,
                        d_frame, d_ping,  d_moments, d_gbuf,    d_rgba8,  d_photons,   d_sorted, d_meta,
                        d_temp,  d_nvalid, d_counters, d_regen, d_acc64, d_gather_rec, d_bframe, d_bmom, d_bgb,
                        d_views, d_tiles, d_counter, d_w4nodes, d_wave, d_relmse, d_display, d_albedo, d_mom_hist, d_fstate,
                        d_tex[0], d_tex[1], d_tex[2], d_tex[3], d_keys, d_surface};
        for (void *b : bufs)
            if (b) hipFree(b);
        if (stream) hipStreamDestroy(stream);
    }

    void begin(int pass, hipEvent_t &start, hipStream_t s = nullptr) {
        start = nullptr;
        if (!(cfg.flags & RT_FLAG_PROFILE)) return;
        hip_check(hipEventCreate(&start), "hipEventCreate");
        hip_check(hipEventRecord(start, s ? s : stream), "hipEventRecord");
        (void)pass;
    }
    void end(int pass, hipEvent_t start, hipStream_t s = nullptr) {
        if (!start) return;
        hipEvent_t stop;
        hip_check(hipEventCreate(&stop), "hipEventCreate");
        hip_check(hipEventRecord(stop, s ? s : stream), "hipEventRecord");
        pending.push_back({pass, start, stop});
    }
    void collect() {
        for (auto &p : pending) {
            hip_check(hipEventSynchronize(p.stop), "hipEventSynchronize");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, p.start, p.stop), "hipEventElapsedTime");
            stats.ms[p.pass] += ms;
            stats.launches[p.pass] += 1;
            hipEventDestroy(p.start);
            hipEventDestroy(p.stop);
        }
        pending.clear();
    }

    rtk::TraceInputs trace_inputs() const {
        rtk::TraceInputs t;
        t.nodes = d_nodes;
        t.prims = d_prims;
        t.slot_tris = traversal_mode() == 4u ? d_slot_tris_wt : d_slot_tris;
        t.wnodes = d_wnodes;
        t.w4nodes = d_w4nodes;
        t.n_nodes = n_nodes;
        std::memcpy(t.root_min, root_min, sizeof(root_min));
        std::memcpy(t.root_max, root_max, sizeof(root_max));
        return t;
    }
    // 0 reference BVH2 nodes, 1 binary child-box records, 2 binary nearest first,
    // 3 4-wide records, 4 4-wide records with the watertight triangle test
    uint32_t traversal_mode() const {
        if (!wide_ok || (cfg.flags & RT_FLAG_REFERENCE_TRAVERSAL)) return 0u;
        if (cfg.flags & RT_FLAG_NEAREST_FIRST) return 2u;
        if ((cfg.flags & RT_FLAG_BINARY_RECORDS) || !wide4_ok) return 1u;
        if ((cfg.flags & RT_FLAG_WATERTIGHT) && d_slot_tris_wt) return 4u;
        return 3u;
    }
    // The watertight traversal's slot triangles, built at the first config
    // that asks for it.
    void ensure_wt_slots() {
        if (d_slot_tris_wt || !(cfg.flags & RT_FLAG_WATERTIGHT) || !wide4_ok) return;
        hip_check(hipMalloc(&d_slot_tris_wt, (size_t)std::max(n_prims, 1u) * 48), "hipMalloc watertight slots");
        hip_check(rtk::launch_wt_slots(d_tris, d_prims, n_prims, d_slot_tris_wt, stream), "watertight slots");
        hip_check(hipStreamSynchronize(stream), "sync");
    }

    rt_image_params image_params(float alpha) const {
        rt_image_params ip;
        std::memset(&ip, 0, sizeof(ip));
        ip.width = cfg.width;
        ip.height = cfg.height;
        ip.spp = cfg.spp;
        ip.frame = sample_counter;
        std::memcpy(ip.light_emission, cfg.light_emission, sizeof(ip.light_emission));
        ip.ambient = cfg.ambient;
        ip.accumulate_alpha = alpha;
        ip.use_photons = cfg.enable_photon_map ? 1u : 0u;
        ip.photon_count = cfg.photon_count;
        ip.photon_radius = cfg.photon_radius;
        ip.photon_scale = cfg.photon_scale;
        ip.photon_debug = cfg.photon_debug ? 1u : 0u;
        std::memcpy(ip.point_light_emission, cfg.point_light_emission, sizeof(ip.point_light_emission));
        ip._pad[0] = cfg.debug_view;
        return ip;
    }

    rtk::GatherLaunch gather_launch(float alpha, float *frame, float *moments, float *gbuf, uint32_t set) const {
        rtk::GatherLaunch g;
        std::memset(&g, 0, sizeof(g));  // padding too: captured launches are compared bytewise
        g.trace = trace_inputs();
        g.tris = d_tris;
        g.mats = d_mats;
        g.lights = d_lights;
        g.n_lights = n_lights;
        g.n_tris = std::max(n_tris, 1u);
        g.cam = cam;
        g.ip = image_params(alpha);
        g.gp = gp;
        g.grid_meta = d_meta;
        g.sorted_photons = d_sorted;
        g.gather_rec = d_gather_rec;
        g.frame = frame;
        g.moments = moments;
        g.gbuffer = gbuf;
        g.max_depth = cfg.max_depth;
        g.mode = traversal_mode();
        // alpha-cutout shadows need the closest hit (gather.wgsl:91-103)
        g.early_shadow = ((cfg.flags & RT_FLAG_CLOSEST_SHADOW) || has_cutout) ? 0u : 1u;
        g.counters = nullptr;
        for (uint32_t k = 0; k < RT_TEX_ROLES; k++) g.tex[k] = tex[k];
        g.px_state = nullptr;
        g.queue[0] = g.queue[1] = nullptr;
        g.path_count = nullptr;
        if (use_path_queues() && d_wave) {
            char *w = static_cast<char *>(d_wave);
            const size_t plane = n_px() * 16u;
            g.px_state = w;
            g.queue[1] = w + 5u * plane;  // paths entering odd depths
            g.queue[0] = w + 8u * plane;  // paths entering even depths
            g.path_count = reinterpret_cast<uint32_t *>(w + 11u * plane);
        }
        // measured on the 1080p workload: all resident workgroups, refill as
        // soon as one lane is free (thresholds 8..32 and 3/4 or 1/2 of the
        // workgroups were no faster)
        g.regen = (cfg.flags & RT_FLAG_TILE_LAUNCH) ? nullptr : d_regen + 8u * set;
        g.regen_blocks = 0u;
        g.refill_min = 1u;
        return g;
    }

    // RT_FLAG_PATH_QUEUES runs the camera pass as path-queue passes; the queue
    // state is sized for the current image.
    bool use_path_queues() const { return (cfg.flags & RT_FLAG_PATH_QUEUES) != 0; }
    void ensure_path_queues() {
        if (!use_path_queues()) return;
        const size_t need = rtk::gather_wavefront_bytes((uint32_t)n_px(), cfg.max_depth);
        if (need <= wave_cap && d_wave) return;
        if (d_wave) hipFree(d_wave);
        d_wave = nullptr;
        wave_cap = 0;
        hip_check(hipMalloc(&d_wave, need), "hipMalloc path queues");
        wave_cap = need;
    }

    float current_alpha() const {
        if (cfg.accumulate_frames == 0) return 1.0f;
        return 1.0f / (float)std::min(frames_since_reset, cfg.accumulate_frames);
    }

    bool track_moments() const { return unet != nullptr || (cfg.flags & RT_FLAG_MOMENT_HISTORY) != 0; }
    // The moment history, allocated zeroed (0 * history is then 0 at a frame
    // with alpha = 1); before any frame graph that records its update.
    void ensure_mom_hist() {
        if (d_mom_hist || !track_moments()) return;
        hip_check(hipMalloc(&d_mom_hist, n_px() * 8), "hipMalloc moment history");
        hip_check(hipMemset(d_mom_hist, 0, n_px() * 8), "clear moment history");
    }

    // Renderer::record_frame, lib.rs:666-851, up to the present pass (rt_present)
    void render_frame() {
        ensure_mom_hist();
        sample_counter = sample_counter + 1u;
        frames_since_reset = frames_since_reset == 0xFFFFFFFFu ? frames_since_reset : frames_since_reset + 1u;
        const float alpha = current_alpha();
        if (cfg.enable_photon_map && !photons_built)
            throw ApiError(RT_ERR_INVALID, "photon map enabled but rt_rebuild_photons has not run");
        ensure_path_queues();
        const uint32_t si = next_set;
        next_set = (next_set + 1u) % (uint32_t)kSets;
        d_frame = set_frame[si];
        d_ping = set_ping[si];
        d_moments = set_mom[si];
        d_gbuf = set_gb[si];
        // the path-queue state is shared by all frames: those gathers stay on the main stream
        hipStream_t gs = use_path_queues() ? stream : gstream[si % kStreams];
        if (gs != stream && free_recorded[si]) hip_check(hipStreamWaitEvent(gs, ev_free[si], 0), "wait set");
        const bool graphs = (cfg.flags & RT_FLAG_GRAPH) != 0;
        rtk::GatherLaunch gl = gather_launch(alpha, d_frame, d_moments, d_gbuf, si);
        hipEvent_t ev;
        if (graphs) {
            ensure_fstate();
            gl.fstate = d_fstate + si;
            ensure_gather_graph(si, gl);
            rtk::FrameState fsv;
            std::memset(&fsv, 0, sizeof(fsv));
            fsv.cam = cam;
            fsv.frame = sample_counter;
            fsv.alpha = alpha;
            hip_check(rtk::launch_set_frame_state(d_fstate + si, fsv, gs), "frame state");
            begin(RT_PASS_GATHER, ev, gs);
            hip_check(hipGraphLaunch(gexec[si], gs), "gather graph");
            end(RT_PASS_GATHER, ev, gs);
        } else {
            begin(RT_PASS_GATHER, ev, gs);
            hip_check(rtk::launch_gather(gl, gs), "gather");
            end(RT_PASS_GATHER, ev, gs);
        }
        if (gs != stream) {
            hip_check(hipEventRecord(ev_gdone[si], gs), "record gather");
            hip_check(hipStreamWaitEvent(stream, ev_gdone[si], 0), "wait gather");
        }
        if (graphs && !unet) {
            ensure_post_graph(si);
            begin(RT_PASS_COMPOSITE_TONEMAP, ev);  // a-trous passes included
            hip_check(hipGraphLaunch(pexec[si], stream), "post graph");
            end(RT_PASS_COMPOSITE_TONEMAP, ev);
        } else {
            post_passes(alpha, nullptr, stream, true);
        }
        post_alpha = alpha;  // a replayed post graph runs no host code
        if (unet) {
            begin(RT_PASS_NETWORK, ev);
            ensure_albedo();
            rtu::denoise(unet, frame_views(), d_display, stream);
            hip_check(rtk::launch_tonemap(d_display, d_rgba8, cfg.width, cfg.height, tonemap_uniform(), d_gbuf, stream),
                      "tonemap");
            end(RT_PASS_NETWORK, ev);
        }
        hip_check(hipEventRecord(ev_free[si], stream), "record set release");
        free_recorded[si] = true;
    }

    // How many of the next n frames rt_render_frames renders with one camera
    // pass launch: up to kChunk with the persistent camera pass (no frame
    // graphs), as long as their output sets are consecutive (no wrap of the
    // set rotation inside a launch); 1 = a frame on its own.
    uint32_t chunk_frames(uint32_t n) const {
        if ((cfg.flags & (RT_FLAG_GRAPH | RT_FLAG_TILE_LAUNCH)) != 0 || use_path_queues()) return 1u;
        return std::max(1u, std::min({n, (uint32_t)kChunk, (uint32_t)kSets - next_set}));
    }

    // nf (2..kChunk) consecutive frames with one camera pass launch: the same
    // frames, seeds, output sets and post passes as nf render_frame calls (the
    // persistent kernel's band counters run over the chunk's frames one after
    // another), so the end of one frame's camera pass, where few long paths
    // are left, overlaps the start of the next frame's.
    void render_chunk(uint32_t nf) {
        ensure_mom_hist();
        if (cfg.enable_photon_map && !photons_built)
            throw ApiError(RT_ERR_INVALID, "photon map enabled but rt_rebuild_photons has not run");
        if (nf < 2u || nf > (uint32_t)kChunk || next_set + nf > (uint32_t)kSets)
            throw ApiError(RT_ERR_INVALID, "frame chunk out of range");
        const uint32_t s0 = next_set;
        next_set = (next_set + nf) % (uint32_t)kSets;
        float alphas[kChunk];
        uint32_t seed0 = 0;
        for (uint32_t f = 0; f < nf; f++) {
            sample_counter = sample_counter + 1u;
            frames_since_reset = frames_since_reset == 0xFFFFFFFFu ? frames_since_reset : frames_since_reset + 1u;
            alphas[f] = current_alpha();
            if (f == 0) seed0 = sample_counter;
        }
        hipStream_t gs = gstream[chunk_stream];  // consecutive chunks alternate between the gather streams
        chunk_stream = (chunk_stream + 1u) % (uint32_t)kStreams;
        for (uint32_t f = 0; f < nf; f++)
            if (free_recorded[s0 + f]) hip_check(hipStreamWaitEvent(gs, ev_free[s0 + f], 0), "wait set");
        rtk::GatherLaunch gl = gather_launch(alphas[0], set_frame[s0], set_mom[s0], set_gb[s0], s0);
        gl.ip.frame = seed0;
        gl.n_frames = nf;
        gl.frame_stride = (uint32_t)n_px();
        hipEvent_t ev;
        begin(RT_PASS_GATHER, ev, gs);
        hip_check(rtk::launch_gather(gl, gs), "gather");
        end(RT_PASS_GATHER, ev, gs);
        hip_check(hipEventRecord(ev_gdone[s0], gs), "record gather");
        hip_check(hipStreamWaitEvent(stream, ev_gdone[s0], 0), "wait gather");
        for (uint32_t f = 0; f < nf; f++) {
            const uint32_t si = s0 + f;
            d_frame = set_frame[si];
            d_ping = set_ping[si];
            d_moments = set_mom[si];
            d_gbuf = set_gb[si];
            post_passes(alphas[f], nullptr, stream, true);
            if (unet) {
                begin(RT_PASS_NETWORK, ev);
                ensure_albedo();
                rtu::denoise(unet, frame_views(), d_display, stream);
                hip_check(rtk::launch_tonemap(d_display, d_rgba8, cfg.width, cfg.height, tonemap_uniform(), d_gbuf,
                                              stream),
                          "tonemap");
                end(RT_PASS_NETWORK, ev);
            }
            hip_check(hipEventRecord(ev_free[si], stream), "record set release");
            free_recorded[si] = true;
        }
    }

    rt_tonemap_uniform tonemap_uniform() const {
        rt_tonemap_uniform tm;
        std::memset(&tm, 0, sizeof(tm));
        tm.exposure = cfg.exposure;
        tm.gamma = cfg.gamma;
        tm.width = cfg.width;
        tm.glow = light_glow;
        if (marker_on && host_cam_ok)
            rt::project_marker(host_cam, cfg.width, cfg.height, rt::Vec3{marker_pos[0], marker_pos[1], marker_pos[2]},
                               marker_radius, tm.marker);
        return tm;
    }

    // a-trous passes, then composite + tonemap, of the current output set on
    // stream s; with fs, alpha comes from the frame state.
    void post_passes(float alpha, const rtk::FrameState *fs, hipStream_t s, bool timed) {
        hipEvent_t ev = nullptr;
        float *result = d_frame;
        if (cfg.enable_denoise) {
            float *src = d_frame, *dst = d_ping;
            for (uint32_t it = 0; it < cfg.denoise_passes; it++) {
                rtk::DenoiseLaunch d;
                d.in = src;
                d.out = dst;
                d.gbuffer = d_gbuf;
                d.moments = d_moments;
                d.p.width = cfg.width;
                d.p.height = cfg.height;
                d.p.step_size = 1u << it;
                d.p.lambda = 3.0f;
                d.p.phi_normal = 12.0f;
                d.p.phi_depth = 0.05f;
                d.p.alpha = alpha;
                d.p._pad = 0;
                d.fstate = fs;
                if (timed) begin(RT_PASS_DENOISE, ev, s);
                hip_check(rtk::launch_denoise(d, s), "denoise");
                if (timed) end(RT_PASS_DENOISE, ev, s);
                std::swap(src, dst);
            }
            result = src;
        }
        if (timed) begin(RT_PASS_COMPOSITE_TONEMAP, ev, s);
        hip_check(rtk::launch_composite_tonemap(result, d_accum, d_rgba8, cfg.width, cfg.height, alpha,
                                                tonemap_uniform(), d_gbuf, fs, s),
                  "composite_tonemap");
        if (track_moments() && d_mom_hist)
            hip_check(rtk::launch_moment_history(d_moments, d_mom_hist, (uint32_t)n_px(), alpha, fs, s),
                      "moment history");
        if (timed) end(RT_PASS_COMPOSITE_TONEMAP, ev, s);
        post_alpha = alpha;
    }

    void sync_streams() {
        hip_check(hipStreamSynchronize(stream), "sync");
        for (int k = 0; k < kStreams; k++) hip_check(hipStreamSynchronize(gstream[k]), "sync");
    }

    // Records f(s) on the capture stream into an instantiated graph.
    template <class F>
    hipGraphExec_t capture(F &&f) {
        if (!cstream) hip_check(hipStreamCreateWithFlags(&cstream, hipStreamNonBlocking), "capture stream");
        hip_check(hipStreamBeginCapture(cstream, hipStreamCaptureModeRelaxed), "begin capture");
        hipGraph_t g = nullptr;
        try {
            f(cstream);
        } catch (...) {
            (void)hipStreamEndCapture(cstream, &g);
            if (g) hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(cstream, &g), "end capture");
        hipGraphExec_t e = nullptr;
        const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        hip_check(err, "instantiate graph");
        graph_captures++;
        return e;
    }

    void ensure_fstate() {
        if (!d_fstate) {
            hip_check(hipMalloc(&d_fstate, sizeof(rtk::FrameState) * kSets), "hipMalloc frame state");
            hip_check(hipMemset(d_fstate, 0, sizeof(rtk::FrameState) * kSets), "memset");
        }
    }

    // The gather graph of set si for launch gl (per-frame fields are ignored:
    // they come from the frame state).
    void ensure_gather_graph(uint32_t si, const rtk::GatherLaunch &gl) {
        rtk::GatherLaunch key = gl;
        std::memset(&key.cam, 0, sizeof(key.cam));
        key.ip.frame = 0;
        key.ip.accumulate_alpha = 0.0f;
        if (gexec[si] && std::memcmp(&key, &gkey[si], sizeof(key)) == 0) return;
        sync_streams();
        if (gexec[si]) hip_check(hipGraphExecDestroy(gexec[si]), "destroy graph");
        gexec[si] = nullptr;
        gexec[si] = capture([&](hipStream_t s) { hip_check(rtk::launch_gather(gl, s), "gather"); });
        gkey[si] = key;
    }

    void ensure_post_graph(uint32_t si) {
        PostKey key;
        std::memset(&key, 0, sizeof(key));
        key.frame = d_frame;
        key.ping = d_ping;
        key.moments = d_moments;
        key.gbuf = d_gbuf;
        key.accum = d_accum;
        key.rgba8 = d_rgba8;
        key.width = cfg.width;
        key.height = cfg.height;
        key.denoise = cfg.enable_denoise;
        key.passes = cfg.denoise_passes;
        key.tm = tonemap_uniform();
        if (pexec[si] && std::memcmp(&key, &pkey[si], sizeof(key)) == 0) return;
        sync_streams();
        if (pexec[si]) hip_check(hipGraphExecDestroy(pexec[si]), "destroy graph");
        pexec[si] = nullptr;
        pexec[si] = capture([&](hipStream_t s) { post_passes(0.0f, d_fstate + si, s, false); });
        pkey[si] = key;
    }

    // First-hit albedo of the pixel-centre rays, recomputed after a camera change.
    void ensure_albedo() {
        if (!d_albedo) {
            hip_check(hipMalloc(&d_albedo, n_px() * 16), "hipMalloc albedo");
            albedo_ok = false;
        }
        if (albedo_ok) return;
        hip_check(rtk::launch_albedo(trace_inputs(), traversal_mode(), d_tris, d_mats, cam, cfg.width, cfg.height,
                                     d_albedo, stream),
                  "albedo");
        albedo_ok = true;
    }
    rtu::FrameViews frame_views() const {
        rtu::FrameViews v;
        v.accum = d_accum;
        v.gbuf = d_gbuf;
        v.albedo = d_albedo;
        v.moments = d_mom_hist;
        v.width = cfg.width;
        v.height = cfg.height;
        v.frames = std::min(frames_since_reset, cfg.accumulate_frames ? cfg.accumulate_frames : 1u);
        v.ema = post_alpha / std::max(2.0f - post_alpha, 1e-4f);
        return v;
    }

    uint32_t traced_photons() const {
        const uint32_t n = std::max(cfg.photon_count, 1u);
        return n >= 256u ? (n / 256u) * 256u : n;
    }
    uint32_t processed_slots() const {
        const uint64_t s = (uint64_t)std::max(cfg.photon_count, 1u) * cfg.max_depth;
        const uint64_t d = std::max<uint64_t>(s / 256u, 1u) * 256u;
        return (uint32_t)std::min<uint64_t>(d, photon_slots);
    }

    void grid_build(uint32_t n_slots, const uint32_t *keys_in) {
        rtk::GridBuildLaunch g;
        g.keys_in = keys_in;
        g.photons = d_photons;
        g.n_slots = n_slots;
        g.gp = gp;
        g.meta_cells = meta_cells;
        g.grid_meta = d_meta;
        g.sorted = d_sorted;
        g.gather_rec = d_gather_rec;
        g.temp = d_temp;
        g.temp_bytes = temp_bytes;
        g.n_valid_out = d_nvalid;
        hipEvent_t ev;
        begin(RT_PASS_GRID_BUILD, ev);
        hip_check(rtk::launch_grid_build(g, stream), "grid build");
        end(RT_PASS_GRID_BUILD, ev);
        hip_check(hipMemcpyAsync(&n_sorted, d_nvalid, 4, hipMemcpyDeviceToHost, stream), "copy n_valid");
        hip_check(hipStreamSynchronize(stream), "sync");
        photons_built = true;
    }

    // Renderer::rebuild_photons, lib.rs:913-1037, with the scan and the
    // scatter on the GPU and the buffers cleared first.
    void rebuild_photons() {
        if (!cfg.enable_photon_map) throw ApiError(RT_ERR_INVALID, "photon map disabled in the config");
        hip_check(hipMemsetAsync(d_photons, 0, (size_t)photon_slots * sizeof(rt_photon), stream), "clear photons");
        hip_check(rtk::launch_fill_u32(d_keys, gp.num_cells, photon_slots, stream), "clear photon keys");
        rtk::PhotonLaunch p;
        p.trace = trace_inputs();
        p.tris = d_tris;
        p.mats = d_mats;
        p.lights = d_lights;
        p.n_lights = n_lights;
        p.n_tris = std::max(n_tris, 1u);
        p.gp = gp;
        p.n_photons = traced_photons();
        p.max_depth = cfg.max_depth;
        for (uint32_t k = 0; k < RT_TEX_ROLES; k++) p.tex[k] = tex[k];
        p.mode = traversal_mode();
        p.photons = d_photons;
        p.keys = d_keys;
        hipEvent_t ev;
        begin(RT_PASS_PHOTON_TRACE, ev);
        hip_check(rtk::launch_photon_trace(p, stream), "photon trace");
        end(RT_PASS_PHOTON_TRACE, ev);
        grid_build(processed_slots(), d_keys);
    }
};

namespace rt {

// Renderer::new grid sizing, lib.rs:416-457; free cell size when requested.
int grid_params_for(const rt_config &cfg, const float *bmin_in, const float *bmax_in, bool has_bounds,
                    rt_grid_params &gp, float &radius, uint32_t &meta_cells) {
    Vec3 bmin{-1.05f, -1.05f, -1.05f}, bmax{1.05f, 1.05f, 1.05f};
    if (has_bounds) {
        bmin = Vec3{bmin_in[0], bmin_in[1], bmin_in[2]};
        bmax = Vec3{bmax_in[0], bmax_in[1], bmax_in[2]};
    }
    const Vec3 ext = bmax - bmin;
    std::memset(&gp, 0, sizeof(gp));
    auto sat_u32 = [](float v) -> uint64_t {
        if (!(v > 0.0f)) return 0;
        if (v >= 4294967296.0f) return 0xFFFFFFFFull;
        return (uint64_t)v;
    };
    float cell;
    uint64_t gx, gy, gz;
    if (cfg.grid_cell_size > 0.0f) {
        cell = cfg.grid_cell_size;
        auto dim = [&](float e) { return std::max<uint64_t>(1, sat_u32(std::floor(e / cell + 0.5f)) + 1); };
        gx = dim(ext.x);
        gy = dim(ext.y);
        gz = dim(ext.z);
        radius = cell;
    } else {
        const float max_ext = std::fmax(max_element(ext), 1e-3f);
        cell = max_ext / ((float)RT_GRID_MAX_DIM - 2.0f);
        auto dim = [&](float e) {
            return std::min<uint64_t>(std::max<uint64_t>(sat_u32(std::ceil(e / cell)), 1), RT_GRID_MAX_DIM);
        };
        gx = dim(ext.x);
        gy = dim(ext.y);
        gz = dim(ext.z);
        radius = std::fmin(std::fmax(cell * 4.0f, 0.01f), 2.0f);
    }
    const uint64_t cells = gx * gy * gz;
    if (cells > (1ull << 28)) return RT_ERR_CAPACITY;
    const Vec3 room_min = bmin - Vec3{cell * 0.5f, cell * 0.5f, cell * 0.5f};
    gp.room_min_x = room_min.x;
    gp.room_min_y = room_min.y;
    gp.room_min_z = room_min.z;
    gp.cell_size = cell;
    gp.grid_x = (uint32_t)gx;
    gp.grid_y = (uint32_t)gy;
    gp.grid_z = (uint32_t)gz;
    gp.num_cells = (uint32_t)cells;
    const uint64_t slots = cfg.enable_photon_map ? std::max<uint64_t>((uint64_t)cfg.photon_count * cfg.max_depth, 1) : 1;
    if (slots > 0xFFFFFFFFull / 2) return RT_ERR_CAPACITY;
    gp.photon_slots = (uint32_t)slots;
    meta_cells = cfg.grid_cell_size > 0.0f ? (uint32_t)cells : std::max<uint32_t>((uint32_t)cells, RT_GRID_MAX_CELLS);
    return RT_OK;
}

}  // namespace rt

using rt::api_call;

static void validate_config(const rt_config &c) {
    if (c.width == 0 || c.height == 0 || (uint64_t)c.width * c.height > (1ull << 28))
        throw ApiError(RT_ERR_INVALID, "width/height must be non-zero and width*height <= 2^28");
    if (c.max_depth == 0 || c.max_depth > 64) throw ApiError(RT_ERR_INVALID, "max_depth must be in 1..64");
    if (c.denoise_passes > 8) throw ApiError(RT_ERR_INVALID, "denoise_passes must be <= 8");
    if (!(c.gamma != 0.0f)) throw ApiError(RT_ERR_INVALID, "gamma must be non-zero");
}

static void select_device(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw ApiError(RT_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) throw ApiError(RT_ERR_NO_DEVICE, "device ordinal out of range");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        throw ApiError(RT_ERR_NO_DEVICE, std::string("kernels are built for gfx950, device is ") + prop.gcnArchName);
    hip_check(hipSetDevice(device), "hipSetDevice");
}

// Derived traversal layouts: slot-ordered triangles and the child-box records.
static void build_trace_layouts(rt_renderer &r, const rt_triangle *tris, const rt_bvh_node *nodes,
                                const uint32_t *prims) {
    std::vector<float> slot((size_t)std::max(r.n_prims, 1u) * 12, 0.0f);
    for (uint32_t k = 0; k < r.n_prims; k++) {
        if (prims[k] >= r.n_tris) throw ApiError(RT_ERR_INVALID, "BVH prim index out of range");
        const rt_triangle &t = tris[prims[k]];
        float *o = &slot[(size_t)k * 12];
        uint32_t pid = prims[k];
        o[0] = t.v0[0]; o[1] = t.v0[1]; o[2] = t.v0[2]; std::memcpy(&o[3], &pid, 4);
        o[4] = t.v1[0] - t.v0[0]; o[5] = t.v1[1] - t.v0[1]; o[6] = t.v1[2] - t.v0[2]; o[7] = 0.0f;
        o[8] = t.v2[0] - t.v0[0]; o[9] = t.v2[1] - t.v0[1]; o[10] = t.v2[2] - t.v0[2]; o[11] = 0.0f;
    }
    r.d_slot_tris = dev_alloc<float>(slot.size(), "slot triangles");
    hip_check(hipMemcpy(r.d_slot_tris, slot.data(), slot.size() * 4, hipMemcpyHostToDevice), "upload slot triangles");

    // validate the BVH and measure its depth
    r.wide_ok = r.n_nodes > 0;
    bool nested = true;
    std::vector<uint32_t> compact(r.n_nodes, 0xFFFFFFFFu);
    std::vector<uint32_t> order;
    if (r.n_nodes > 0) {
        std::vector<std::pair<uint32_t, uint32_t>> st{{0u, 1u}};
        std::vector<uint8_t> seen(r.n_nodes, 0);
        while (!st.empty()) {
            auto [i, d] = st.back();
            st.pop_back();
            if (i >= r.n_nodes || seen[i]) throw ApiError(RT_ERR_INVALID, "BVH child index invalid or shared");
            seen[i] = 1;
            r.bvh_depth = std::max(r.bvh_depth, d);
            const rt_bvh_node &n = nodes[i];
            if (n.prim_count > 0) {
                if ((uint64_t)n.left_first + n.prim_count > r.n_prims)
                    throw ApiError(RT_ERR_INVALID, "BVH leaf range out of range");
                if (n.prim_count > 16 || n.left_first >= (1u << 27)) r.wide_ok = false;
            } else {
                // the 4-wide records skip intermediate boxes: they need children inside parents
                for (uint32_t c : {n.left_first, n.right_child}) {
                    if (c >= r.n_nodes) continue;  // reported when visited
                    for (int a = 0; a < 3; a++)
                        if (!(nodes[c].bbox_min[a] >= n.bbox_min[a] && nodes[c].bbox_max[a] <= n.bbox_max[a]))
                            nested = false;
                }
                compact[i] = (uint32_t)order.size();
                order.push_back(i);
                st.push_back({n.right_child, d + 1});
                st.push_back({n.left_first, d + 1});
            }
        }
    }
    if (r.bvh_depth > 63) r.wide_ok = false;  // the reference would drop subtrees; keep its traversal
    r.n_internal = (uint32_t)order.size();
    auto ref_of = [&](uint32_t c) -> uint32_t {
        const rt_bvh_node &n = nodes[c];
        if (n.prim_count > 0) return 0x80000000u | ((n.prim_count - 1u) << 27) | n.left_first;
        return compact[c];
    };
    std::vector<float> w((size_t)std::max(r.n_internal, 1u) * 16, 0.0f);
    if (r.wide_ok) {
        for (uint32_t k = 0; k < r.n_internal; k++) {
            const rt_bvh_node &n = nodes[order[k]];
            const rt_bvh_node &L = nodes[n.left_first], &R = nodes[n.right_child];
            float *o = &w[(size_t)k * 16];
            const uint32_t lref = ref_of(n.left_first), rref = ref_of(n.right_child);
            o[0] = L.bbox_min[0]; o[1] = L.bbox_min[1]; o[2] = L.bbox_min[2]; std::memcpy(&o[3], &lref, 4);
            o[4] = L.bbox_max[0]; o[5] = L.bbox_max[1]; o[6] = L.bbox_max[2]; std::memcpy(&o[7], &rref, 4);
            o[8] = R.bbox_min[0]; o[9] = R.bbox_min[1]; o[10] = R.bbox_min[2];
            o[12] = R.bbox_max[0]; o[13] = R.bbox_max[1]; o[14] = R.bbox_max[2];
        }
        const rt_bvh_node &root = nodes[0];
        const uint32_t rootref = ref_of(0);
        r.root_min[0] = root.bbox_min[0]; r.root_min[1] = root.bbox_min[1]; r.root_min[2] = root.bbox_min[2];
        std::memcpy(&r.root_min[3], &rootref, 4);
        r.root_max[0] = root.bbox_max[0]; r.root_max[1] = root.bbox_max[1]; r.root_max[2] = root.bbox_max[2];
    }
    r.d_wnodes = dev_alloc<float>(w.size(), "wide nodes");
    hip_check(hipMemcpy(r.d_wnodes, w.data(), w.size() * 4, hipMemcpyHostToDevice), "upload wide nodes");
    // Stack words the traversals can need: one deferred (ref, te) pair per level
    // for the binary records, (children - 1) pairs per level for the 4-wide ones.
    if (2u * (r.bvh_depth > 0 ? r.bvh_depth - 1 : 0) > kTraversalStackWords) r.wide_ok = false;

    // 4-wide records: each internal node's children are expanded (largest box
    // first) into up to 4 descendants, keeping their left-to-right order, so a
    // left-to-right depth-first walk meets the leaves in the reference's order.
    // Record: min x/y/z, max x/y/z of the 4 children (SoA, 6 float4), refs (uint4),
    // kill bits, 12 bytes of padding; absent children have ref 0xFFFFFFFF.
    r.wide4_ok = r.wide_ok && nested;
    std::vector<float> w4;
    if (r.wide4_ok) {
        auto area = [&](uint32_t i) {
            const rt_bvh_node &n = nodes[i];
            const float ex = n.bbox_max[0] - n.bbox_min[0], ey = n.bbox_max[1] - n.bbox_min[1],
                        ez = n.bbox_max[2] - n.bbox_min[2];
            return ex * ey + ey * ez + ez * ex;
        };
        // records in creation order (parents before children); kids: BVH2 node indices
        std::vector<uint32_t> src;                   // BVH2 node of each record
        std::vector<std::vector<uint32_t>> kids_of;  // BVH2 children of each record
        std::vector<std::array<uint32_t, 4>> refs_of;
        std::vector<uint32_t> kill_of;               // 3 kill bits per child
        if (nodes[0].prim_count == 0) src.push_back(0u);
        for (size_t rec = 0; rec < src.size(); rec++) {
            std::vector<uint32_t> kids{nodes[src[rec]].left_first, nodes[src[rec]].right_child};
            std::vector<std::vector<uint32_t>> skipped(2);  // untested BVH2 nodes above each child
            while (kids.size() < 4) {
                int best = -1;
                float best_a = -1.0f;
                for (size_t k = 0; k < kids.size(); k++)
                    if (nodes[kids[k]].prim_count == 0 && area(kids[k]) > best_a) {
                        best_a = area(kids[k]);
                        best = (int)k;
                    }
                if (best < 0) break;
                const uint32_t c = kids[best];
                std::vector<uint32_t> sk = skipped[best];
                sk.push_back(c);
                kids[best] = nodes[c].left_first;
                kids.insert(kids.begin() + best + 1, nodes[c].right_child);
                skipped[best] = sk;
                skipped.insert(skipped.begin() + best + 1, sk);
            }
            // kill bit (child k, axis a): the child is flat on a at a face of a
            // skipped node that is not flat on a
            uint32_t kill = 0;
            for (size_t k = 0; k < kids.size(); k++) {
                const rt_bvh_node &c = nodes[kids[k]];
                for (int a = 0; a < 3; a++) {
                    if (c.bbox_min[a] != c.bbox_max[a]) continue;
                    for (uint32_t m : skipped[k]) {
                        const rt_bvh_node &M = nodes[m];
                        if (M.bbox_min[a] != M.bbox_max[a] &&
                            (c.bbox_min[a] == M.bbox_min[a] || c.bbox_min[a] == M.bbox_max[a]))
                            kill |= 1u << (3 * k + a);
                    }
                }
            }
            kill_of.push_back(kill);
            std::array<uint32_t, 4> refs{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
            for (size_t k = 0; k < kids.size(); k++) {
                const rt_bvh_node &c = nodes[kids[k]];
                if (c.prim_count > 0) {
                    refs[k] = 0x80000000u | ((c.prim_count - 1u) << 27) | c.left_first;
                } else {
                    refs[k] = (uint32_t)src.size();
                    src.push_back(kids[k]);
                }
            }
            kids_of.push_back(kids);
            refs_of.push_back(refs);
        }
        const uint32_t n_rec = (uint32_t)src.size();
        // worst-case deferred pairs from each record down (children come after parents)
        std::vector<uint32_t> need(n_rec, 0);
        for (size_t rec = n_rec; rec-- > 0;) {
            uint32_t deeper = 0;
            for (uint32_t ref : refs_of[rec])
                if (ref != 0xFFFFFFFFu && !(ref & 0x80000000u)) deeper = std::max(deeper, need[ref]);
            need[rec] = (uint32_t)kids_of[rec].size() - 1u + deeper;
        }
        r.stack_words4 = n_rec > 0 ? 2u * need[0] : 0u;
        if (n_rec > 0 && 2u * need[0] > kTraversalStackWords) r.wide4_ok = false;
        w4.assign((size_t)std::max(n_rec, 1u) * 32, 0.0f);
        for (uint32_t rec = 0; rec < n_rec && r.wide4_ok; rec++) {
            float *o = &w4[(size_t)rec * 32];
            for (size_t k = 0; k < kids_of[rec].size(); k++) {
                const rt_bvh_node &c = nodes[kids_of[rec][k]];
                for (int a = 0; a < 3; a++) {
                    o[a * 4 + k] = c.bbox_min[a];
                    o[12 + a * 4 + k] = c.bbox_max[a];
                }
            }
            std::memcpy(&o[24], refs_of[rec].data(), 16);
            std::memcpy(&o[28], &kill_of[rec], 4);
        }
        r.n_rec4 = n_rec;
    }
    if (w4.empty()) w4.assign(32, 0.0f);
    r.d_w4nodes = dev_alloc<float>(w4.size(), "4-wide nodes");
    hip_check(hipMemcpy(r.d_w4nodes, w4.data(), w4.size() * 4, hipMemcpyHostToDevice), "upload 4-wide nodes");
}

static void upload_texture_array(rt_renderer &r, uint32_t role, const rt::TextureArray &a) {
    if (a.width == 0 || a.height == 0 || a.layers == 0 ||
        a.texels.size() != (size_t)a.width * a.height * a.layers * 4)
        throw ApiError(RT_ERR_INVALID, "texture array size mismatch");
    uint8_t *d = dev_alloc<uint8_t>(a.texels.size(), "texture array");
    hip_check(hipMemcpy(d, a.texels.data(), a.texels.size(), hipMemcpyHostToDevice), "upload texture array");
    if (r.d_tex[role]) hip_check(hipFree(r.d_tex[role]), "free texture array");
    r.d_tex[role] = d;
    r.tex[role] = rt_tex_array{d, a.width, a.height, a.layers, 0u};
}

static rt_renderer *create_renderer(const rt_config &cfg_in, const rt_triangle *tris, uint32_t n_tris,
                                    const rt_bvh_node *nodes, uint32_t n_nodes, const uint32_t *prims, uint32_t n_prims,
                                    const rt_material *mats, uint32_t n_mats, const rt_light *lights, uint32_t n_lights,
                                    const float *bmin, const float *bmax, int device,
                                    const rt::TextureArray *textures) {
    validate_config(cfg_in);
    for (uint32_t i = 0; i < n_tris; i++)
        if (tris[i].material_id >= std::max(n_mats, 1u))
            throw ApiError(RT_ERR_INVALID, "triangle material id out of range");
    if ((n_nodes == 0) != (n_prims == 0) && n_tris > 0) throw ApiError(RT_ERR_INVALID, "inconsistent BVH arrays");
    select_device(device);
    auto r = std::make_unique<rt_renderer>();
    r->cfg = cfg_in;
    r->device = device;
    hip_check(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking), "hipStreamCreate");
    r->n_tris = n_tris;
    r->n_nodes = n_nodes;
    r->n_prims = n_prims;
    r->n_mats = n_mats;
    r->n_lights = std::max(n_lights, 1u);
    for (uint32_t i = 0; i < n_mats; i++)
        if ((mats[i].flags & RT_MAT_FLAG_ALPHA_CUTOUT) && mats[i].tex_albedo != RT_TEX_NONE) r->has_cutout = true;

    // scene upload (lib.rs:234-271); empty arrays bind one zeroed record (buffer.rs:19-34)
    r->d_tris = dev_alloc<rt_triangle>(n_tris, "triangles");
    r->d_nodes = dev_alloc<rt_bvh_node>(n_nodes, "bvh nodes");
    r->d_prims = dev_alloc<uint32_t>(n_prims, "bvh prims");
    r->d_mats = dev_alloc<rt_material>(n_mats, "materials");
    r->d_lights = dev_alloc<rt_light>(r->n_lights, "lights");
    hip_check(hipMemset(r->d_tris, 0, sizeof(rt_triangle) * std::max(n_tris, 1u)), "memset");
    hip_check(hipMemset(r->d_mats, 0, sizeof(rt_material) * std::max(n_mats, 1u)), "memset");
    hip_check(hipMemset(r->d_lights, 0, sizeof(rt_light) * r->n_lights), "memset");
    if (n_tris) hip_check(hipMemcpy(r->d_tris, tris, sizeof(rt_triangle) * n_tris, hipMemcpyHostToDevice), "upload");
    if (n_nodes) hip_check(hipMemcpy(r->d_nodes, nodes, sizeof(rt_bvh_node) * n_nodes, hipMemcpyHostToDevice), "upload");
    if (n_prims) hip_check(hipMemcpy(r->d_prims, prims, 4ull * n_prims, hipMemcpyHostToDevice), "upload");
    if (n_mats) hip_check(hipMemcpy(r->d_mats, mats, sizeof(rt_material) * n_mats, hipMemcpyHostToDevice), "upload");
    if (n_lights) hip_check(hipMemcpy(r->d_lights, lights, sizeof(rt_light) * n_lights, hipMemcpyHostToDevice), "upload");
    build_trace_layouts(*r, tris, nodes, prims);

    // per-pixel buffers (lib.rs:287-316)
    const size_t npx = r->n_px();
    r->d_accum = dev_alloc<float>(npx * 4, "accumulation");
    const size_t nset = (size_t)rt_renderer::kSets;
    r->blk_frame = dev_alloc<float>(nset * npx * 4, "frame sets");
    r->blk_ping = dev_alloc<float>(nset * npx * 4, "denoise pingpong sets");
    r->blk_mom = dev_alloc<float>(nset * npx * 4, "frame moment sets");
    r->blk_gb = dev_alloc<float>(nset * npx * 4, "gbuffer sets");
    r->d_rgba8 = dev_alloc<uint32_t>(npx, "output");
    for (int k = 0; k < rt_renderer::kSets; k++) {
        r->set_frame[k] = r->blk_frame + (size_t)k * npx * 4;
        r->set_ping[k] = r->blk_ping + (size_t)k * npx * 4;
        r->set_mom[k] = r->blk_mom + (size_t)k * npx * 4;
        r->set_gb[k] = r->blk_gb + (size_t)k * npx * 4;
    }
    r->d_frame = r->set_frame[0];
    r->d_ping = r->set_ping[0];
    r->d_moments = r->set_mom[0];
    r->d_gbuf = r->set_gb[0];
    hip_check(hipMemset(r->d_accum, 0, npx * 16), "memset");
    for (float *b : {r->blk_frame, r->blk_ping, r->blk_mom, r->blk_gb})
        hip_check(hipMemset(b, 0, nset * npx * 16), "memset");
    for (int k = 0; k < rt_renderer::kStreams; k++)
        hip_check(hipStreamCreateWithFlags(&r->gstream[k], hipStreamNonBlocking), "hipStreamCreate");
    for (int k = 0; k < rt_renderer::kSets; k++) {
        hip_check(hipEventCreateWithFlags(&r->ev_gdone[k], hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&r->ev_free[k], hipEventDisableTiming), "hipEventCreate");
    }
    hip_check(hipMemset(r->d_rgba8, 0, npx * 4), "memset");

    // photon grid (lib.rs:416-479)
    float radius = 0.0f;
    const int gs = rt::grid_params_for(r->cfg, bmin, bmax, bmin && bmax, r->gp, radius, r->meta_cells);
    if (gs != RT_OK) throw ApiError(gs, "photon grid or photon buffer exceeds the supported size");
    r->cfg.photon_radius = radius;
    r->photon_slots = r->gp.photon_slots;
    r->d_photons = dev_alloc<rt_photon>(r->photon_slots, "photons");
    r->d_sorted = dev_alloc<rt_photon>(r->photon_slots, "sorted photons");
    r->d_keys = dev_alloc<uint32_t>(r->photon_slots, "photon keys");
    r->d_gather_rec = dev_alloc<float>((size_t)r->photon_slots * 12, "photon gather records");
    hip_check(hipMemset(r->d_gather_rec, 0, 48ull * r->photon_slots), "memset");
    r->d_meta = dev_alloc<uint32_t>((size_t)r->meta_cells * 2, "grid meta");
    hip_check(hipMemset(r->d_photons, 0, sizeof(rt_photon) * r->photon_slots), "memset");
    hip_check(hipMemset(r->d_sorted, 0, sizeof(rt_photon) * r->photon_slots), "memset");
    hip_check(hipMemset(r->d_meta, 0, 8ull * r->meta_cells), "memset");
    r->temp_bytes = rtk::grid_build_temp_bytes(r->photon_slots, r->meta_cells);
    r->d_temp = dev_alloc<char>(r->temp_bytes, "grid temp");
    r->d_nvalid = dev_alloc<uint32_t>(1, "n_valid");
    r->d_counters = dev_alloc<unsigned long long>(kCounterWords, "counters");
    r->d_regen = dev_alloc<uint32_t>(8u * (rt_renderer::kSets + 1u), "pixel counters");

    // texture arrays: the scene's, else the reference's 1x1 placeholder per role
    for (uint32_t k = 0; k < RT_TEX_ROLES; k++)
        upload_texture_array(*r, k, textures ? textures[k] : rt::TextureArray{});

    // default camera (lib.rs:266-270)
    r->cam = rt::camera_to_uniform(rt::camera_default(), r->cfg.width, r->cfg.height);
    r->ensure_wt_slots();
    hip_check(hipDeviceSynchronize(), "upload sync");
    return r.release();
}

extern "C" {

int rt_config_default(rt_config *c) {
    return api_call([&] {
        if (!c) throw ApiError(RT_ERR_INVALID, "null config");
        std::memset(c, 0, sizeof(*c));
        c->width = 800;
        c->height = 600;
        c->photon_count = 1000000;
        c->spp = 1;
        c->exposure = 1.0f;
        c->gamma = 2.2f;
        for (int k = 0; k < 3; k++) c->light_emission[k] = c->point_light_emission[k] = 1.0f;
        c->accumulate_frames = 0;
        c->enable_photon_map = 0;
        c->enable_denoise = 1;
        c->photon_debug = 0;
        c->debug_view = 0;
        c->photon_scale = 0.02f;
        c->photon_radius = 0.5f;
        c->ambient = 0.03f;
        c->max_depth = RT_MAX_BOUNCES;
        c->grid_cell_size = 0.0f;
        c->denoise_passes = 2;
        c->flags = 0;
    });
}

int rt_grid_params_for(const rt_config *cfg, const float bmin[3], const float bmax[3], int has_bounds,
                       rt_grid_params *out, float *photon_radius) {
    return api_call([&] {
        if (!cfg || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        float radius = 0.0f;
        uint32_t meta = 0;
        const int s = rt::grid_params_for(*cfg, bmin, bmax, has_bounds && bmin && bmax, *out, radius, meta);
        if (s != RT_OK) throw ApiError(s, "grid too large");
        if (photon_radius) *photon_radius = radius;
    });
}

int rt_renderer_create_raw(const rt_config *cfg, const rt_triangle *tris, uint32_t n_tris, const rt_bvh_node *nodes,
                           uint32_t n_nodes, const uint32_t *prims, uint32_t n_prims, const rt_material *mats,
                           uint32_t n_mats, const rt_light *lights, uint32_t n_lights, const float *bounds_min,
                           const float *bounds_max, int device, rt_renderer **out) {
    return api_call([&] {
        if (!cfg || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        if ((n_tris && !tris) || (n_nodes && !nodes) || (n_prims && !prims) || (n_mats && !mats) ||
            (n_lights && !lights))
            throw ApiError(RT_ERR_INVALID, "null array with non-zero count");
        *out = create_renderer(*cfg, tris, n_tris, nodes, n_nodes, prims, n_prims, mats, n_mats, lights, n_lights,
                               bounds_min, bounds_max, device, nullptr);
    });
}

int rt_renderer_set_texture_array(rt_renderer *r, uint32_t role, uint32_t width, uint32_t height, uint32_t layers,
                                  const uint8_t *texels) {
    return api_call([&] {
        if (!r || !texels) throw ApiError(RT_ERR_INVALID, "null argument");
        if (role >= RT_TEX_ROLES) throw ApiError(RT_ERR_INVALID, "texture role out of range");
        if (width == 0 || height == 0 || layers == 0 || (uint64_t)width * height * layers > (1ull << 28))
            throw ApiError(RT_ERR_INVALID, "texture array size out of range");
        hip_check(hipStreamSynchronize(r->stream), "sync");
        rt::TextureArray a;
        a.width = width;
        a.height = height;
        a.layers = layers;
        a.texels.assign(texels, texels + (size_t)width * height * layers * 4);
        upload_texture_array(*r, role, a);
        r->albedo_ok = false;
    });
}

int rt_renderer_create(const rt_config *cfg, const rt_scene *scene, int device, rt_renderer **out) {
    return api_call([&] {
        if (!cfg || !scene || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        const rt::Scene &s = rt::scene_of(scene);
        const std::vector<rt_triangle> tris = s.flatten_triangles();
        const rt::BVH bvh = s.build_bvh();
        const std::vector<rt_light> lights = s.lights_gpu();
        rt::Vec3 bmin, bmax;
        const bool hb = s.bounds(bmin, bmax);
        const float lo[3] = {bmin.x, bmin.y, bmin.z}, hi[3] = {bmax.x, bmax.y, bmax.z};
        rt::TextureArray arrays[RT_TEX_ROLES];
        for (uint32_t k = 0; k < RT_TEX_ROLES; k++) arrays[k] = rt::build_texture_array(s.textures[k]);
        *out = create_renderer(*cfg, tris.data(), (uint32_t)tris.size(), bvh.nodes.data(), (uint32_t)bvh.nodes.size(),
                               bvh.prim_indices.data(), (uint32_t)bvh.prim_indices.size(), s.materials.data(),
                               (uint32_t)s.materials.size(), lights.data(), (uint32_t)lights.size(),
                               hb ? lo : nullptr, hb ? hi : nullptr, device, arrays);
    });
}

void rt_renderer_destroy(rt_renderer *r) { delete r; }

int rt_render_frame(rt_renderer *r) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        r->render_frame();
    });
}

int rt_render_frames(rt_renderer *r, uint32_t n) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        while (n > 0) {
            const uint32_t nf = r->chunk_frames(n);
            if (nf >= 2u) {
                r->render_chunk(nf);
                n -= nf;
            } else {
                r->render_frame();
                n--;
            }
        }
    });
}

int rt_synchronize(rt_renderer *r) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        hip_check(hipStreamSynchronize(r->stream), "hipStreamSynchronize");
        r->collect();
    });
}

int rt_update_camera(rt_renderer *r, const rt_camera *cam) {
    return api_call([&] {
        if (!r || !cam) throw ApiError(RT_ERR_INVALID, "null argument");
        r->cam = rt::camera_to_uniform(*cam, r->cfg.width, r->cfg.height);
        r->host_cam = *cam;
        r->host_cam_ok = true;
        r->albedo_ok = false;
    });
}

int rt_update_camera_uniform(rt_renderer *r, const rt_camera_uniform *u) {
    return api_call([&] {
        if (!r || !u) throw ApiError(RT_ERR_INVALID, "null argument");
        r->cam = *u;
        r->host_cam_ok = false;
        r->albedo_ok = false;
    });
}

int rt_update_lights(rt_renderer *r, const rt_light *lights, uint32_t n) {
    return api_call([&] {
        if (!r || (n && !lights)) throw ApiError(RT_ERR_INVALID, "null argument");
        if (n > r->n_lights) throw ApiError(RT_ERR_CAPACITY, "more lights than uploaded at creation");
        hip_check(hipStreamSynchronize(r->stream), "sync");
        if (n) hip_check(hipMemcpy(r->d_lights, lights, sizeof(rt_light) * n, hipMemcpyHostToDevice), "upload lights");
        // TypedBuffer::write overwrites the leading records only (buffer.rs:64-66)
    });
}

int rt_reset_accumulation(rt_renderer *r) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        hip_check(hipMemsetAsync(r->d_accum, 0, r->n_px() * 16, r->stream), "clear accumulation");
        r->frames_since_reset = 0;
    });
}

int rt_reset_accumulation_soft(rt_renderer *r) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        r->frames_since_reset = 0;
    });
}

int rt_rebuild_photons(rt_renderer *r) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        r->rebuild_photons();
    });
}

uint32_t rt_sample_count(const rt_renderer *r) { return r ? r->sample_counter : 0u; }

int rt_set_sample_count(rt_renderer *r, uint32_t value) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        r->sample_counter = value;
    });
}

int rt_get_config(const rt_renderer *r, rt_config *out) {
    return api_call([&] {
        if (!r || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        *out = r->cfg;
    });
}

int rt_set_config(rt_renderer *r, const rt_config *cfg) {
    return api_call([&] {
        if (!r || !cfg) throw ApiError(RT_ERR_INVALID, "null argument");
        if (cfg->width != r->cfg.width || cfg->height != r->cfg.height || cfg->photon_count != r->cfg.photon_count ||
            cfg->max_depth != r->cfg.max_depth || cfg->grid_cell_size != r->cfg.grid_cell_size ||
            cfg->enable_photon_map != r->cfg.enable_photon_map)
            throw ApiError(RT_ERR_INVALID, "size, photon and grid fields are fixed at creation");
        validate_config(*cfg);
        r->cfg = *cfg;
        r->ensure_wt_slots();
    });
}

int rt_get_grid_params(const rt_renderer *r, rt_grid_params *out) {
    return api_call([&] {
        if (!r || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        *out = r->gp;
    });
}

static void read_dev(rt_renderer *r, void *dst, const void *src, size_t bytes) {
    hip_check(hipStreamSynchronize(r->stream), "sync");
    hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "read back");
}

static int read_px4(rt_renderer *r, float *dst, size_t n, const float *src) {
    return api_call([&] {
        if (!r || !dst) throw ApiError(RT_ERR_INVALID, "null argument");
        if (n < r->n_px() * 4) throw ApiError(RT_ERR_INVALID, "buffer smaller than width*height*4 floats");
        read_dev(r, dst, src, r->n_px() * 16);
    });
}

int rt_read_accumulation(rt_renderer *r, float *rgba, size_t n) { return read_px4(r, rgba, n, r ? r->d_accum : nullptr); }
int rt_read_frame(rt_renderer *r, float *rgba, size_t n) { return read_px4(r, rgba, n, r ? r->d_frame : nullptr); }
int rt_read_moments(rt_renderer *r, float *rgba, size_t n) { return read_px4(r, rgba, n, r ? r->d_moments : nullptr); }
int rt_read_gbuffer(rt_renderer *r, float *rgba, size_t n) { return read_px4(r, rgba, n, r ? r->d_gbuf : nullptr); }

int rt_read_output_rgba8(rt_renderer *r, uint8_t *rgba, size_t n) {
    return api_call([&] {
        if (!r || !rgba) throw ApiError(RT_ERR_INVALID, "null argument");
        if (n < r->n_px() * 4) throw ApiError(RT_ERR_INVALID, "buffer smaller than width*height*4 bytes");
        read_dev(r, rgba, r->d_rgba8, r->n_px() * 4);
    });
}

// Largest surface side the present pass accepts (pixel centres stay exact in float32).
static constexpr uint32_t kMaxSurfaceSide = 32768;

static void present_into(rt_renderer *r, uint32_t *d_surface, uint32_t sw, uint32_t sh, size_t pitch_words) {
    hipEvent_t ev;
    r->begin(RT_PASS_PRESENT, ev);
    hip_check(rtk::launch_present(r->d_rgba8, r->cfg.width, r->cfg.height, d_surface, sw, sh, (uint32_t)pitch_words,
                                  r->stream),
              "present");
    r->end(RT_PASS_PRESENT, ev);
}

static void check_surface(uint32_t sw, uint32_t sh) {
    if (sw == 0 || sh == 0 || sw > kMaxSurfaceSide || sh > kMaxSurfaceSide)
        throw ApiError(RT_ERR_INVALID, "surface size must be 1..32768 in each dimension");
}

int rt_present(rt_renderer *r, void *d_surface, uint32_t surface_w, uint32_t surface_h, size_t row_pitch) {
    return api_call([&] {
        if (!r || !d_surface) throw ApiError(RT_ERR_INVALID, "null argument");
        check_surface(surface_w, surface_h);
        if (row_pitch % 4 != 0 || row_pitch < (size_t)surface_w * 4 || row_pitch / 4 > 0xFFFFFFFFu)
            throw ApiError(RT_ERR_INVALID, "row_pitch must be a multiple of 4 and at least surface_w*4 bytes");
        present_into(r, static_cast<uint32_t *>(d_surface), surface_w, surface_h, row_pitch / 4);
    });
}

int rt_read_present(rt_renderer *r, uint32_t surface_w, uint32_t surface_h, uint8_t *bgra, size_t n_bytes) {
    return api_call([&] {
        if (!r || !bgra) throw ApiError(RT_ERR_INVALID, "null argument");
        check_surface(surface_w, surface_h);
        const size_t words = (size_t)surface_w * surface_h;
        if (n_bytes < words * 4) throw ApiError(RT_ERR_INVALID, "buffer smaller than surface_w*surface_h*4 bytes");
        if (words > r->surface_words) {
            hip_check(hipStreamSynchronize(r->stream), "sync");
            if (r->d_surface) hip_check(hipFree(r->d_surface), "free surface");
            r->d_surface = nullptr;
            r->surface_words = 0;
            r->d_surface = dev_alloc<uint32_t>(words, "surface");
            r->surface_words = words;
        }
        present_into(r, r->d_surface, surface_w, surface_h, surface_w);
        read_dev(r, bgra, r->d_surface, words * 4);
    });
}

int rt_write_accumulation(rt_renderer *r, const float *rgba, size_t n) {
    return api_call([&] {
        if (!r || !rgba) throw ApiError(RT_ERR_INVALID, "null argument");
        if (n < r->n_px() * 4) throw ApiError(RT_ERR_INVALID, "buffer smaller than width*height*4 floats");
        hip_check(hipStreamSynchronize(r->stream), "sync");
        hip_check(hipMemcpy(r->d_accum, rgba, r->n_px() * 16, hipMemcpyHostToDevice), "write accumulation");
    });
}

int rt_read_photons(rt_renderer *r, rt_photon *out, size_t cap, uint32_t *n_slots) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        if (n_slots) *n_slots = r->photon_slots;
        if (!out) return;
        if (cap < r->photon_slots) throw ApiError(RT_ERR_INVALID, "capacity below photon slot count");
        read_dev(r, out, r->d_photons, sizeof(rt_photon) * r->photon_slots);
    });
}

int rt_read_sorted_photons(rt_renderer *r, rt_photon *out, size_t cap, uint32_t *n_stored) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        if (n_stored) *n_stored = r->n_sorted;
        if (!out) return;
        if (cap < r->n_sorted) throw ApiError(RT_ERR_INVALID, "capacity below stored photon count");
        if (r->n_sorted) read_dev(r, out, r->d_sorted, sizeof(rt_photon) * r->n_sorted);
    });
}

int rt_read_grid_meta(rt_renderer *r, uint32_t *meta, size_t n, uint32_t *n_cells) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        if (n_cells) *n_cells = r->meta_cells;
        if (!meta) return;
        if (n < 2ull * r->meta_cells) throw ApiError(RT_ERR_INVALID, "capacity below 2*cells");
        read_dev(r, meta, r->d_meta, 8ull * r->meta_cells);
    });
}

int rt_set_light_marker(rt_renderer *r, const float *position, float radius) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        r->marker_on = position != nullptr;
        if (position) {
            std::memcpy(r->marker_pos, position, sizeof(r->marker_pos));
            r->marker_radius = radius;
        }
    });
}

int rt_set_light_glow(rt_renderer *r, float glow) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        r->light_glow = glow;
    });
}

int rt_project_marker(const rt_camera *cam, uint32_t width, uint32_t height, const float position[3], float radius,
                      float out[4]) {
    return api_call([&] {
        if (!cam || !position || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        if (width == 0 || height == 0) throw ApiError(RT_ERR_INVALID, "zero image size");
        rt::project_marker(*cam, width, height, rt::Vec3{position[0], position[1], position[2]}, radius, out);
    });
}

int rt_frame_graph_captures(const rt_renderer *r, uint32_t *captures) {
    return api_call([&] {
        if (!r || !captures) throw ApiError(RT_ERR_INVALID, "null argument");
        *captures = r->graph_captures;
    });
}

int rt_traversal_info(const rt_renderer *r, uint32_t *mode, uint32_t *bvh_depth, uint32_t *n_records4) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        if (mode) *mode = r->traversal_mode();
        if (bvh_depth) *bvh_depth = r->bvh_depth;
        if (n_records4) *n_records4 = r->wide4_ok ? r->n_rec4 : 0u;
    });
}

int rt_traversal_stack_words(const rt_renderer *r, uint32_t *words) {
    return api_call([&] {
        if (!r || !words) throw ApiError(RT_ERR_INVALID, "null argument");
        // a deferred child is one word (reference nodes, at most 63 pushed) or two
        // (child-box records: index and entry distance); a walk defers at most one
        // child per level below the root
        const uint32_t m = r->traversal_mode();
        const uint32_t deferred = r->bvh_depth > 0 ? r->bvh_depth - 1 : 0;
        *words = m == 3u ? r->stack_words4 : (m == 0u ? std::min(63u, deferred) : 2u * deferred);
    });
}

int rt_trace_closest_device(rt_renderer *r, const rt_ray *d_rays, uint32_t n, rt_hit *d_out) {
    return api_call([&] {
        if (!r || (n && (!d_rays || !d_out))) throw ApiError(RT_ERR_INVALID, "null argument");
        hipEvent_t ev;
        r->begin(RT_PASS_TRACE_RAYS, ev);
        hip_check(rtk::launch_trace_rays(r->trace_inputs(), r->traversal_mode(), d_rays, n, d_out, r->stream),
                  "trace rays");
        r->end(RT_PASS_TRACE_RAYS, ev);
    });
}

int rt_trace_closest(rt_renderer *r, const rt_ray *rays, uint32_t n, rt_hit *out) {
    return api_call([&] {
        if (!r || (n && (!rays || !out))) throw ApiError(RT_ERR_INVALID, "null argument");
        if (n == 0) return;
        rt_ray *d_r = dev_alloc<rt_ray>(n, "rays");
        rt_hit *d_h = nullptr;
        try {
            d_h = dev_alloc<rt_hit>(n, "hits");
            hip_check(hipMemcpy(d_r, rays, sizeof(rt_ray) * n, hipMemcpyHostToDevice), "upload rays");
            hipEvent_t ev;
            r->begin(RT_PASS_TRACE_RAYS, ev);
            hip_check(rtk::launch_trace_rays(r->trace_inputs(), r->traversal_mode(), d_r, n, d_h, r->stream),
                      "trace rays");
            r->end(RT_PASS_TRACE_RAYS, ev);
            hip_check(hipStreamSynchronize(r->stream), "sync");
            hip_check(hipMemcpy(out, d_h, sizeof(rt_hit) * n, hipMemcpyDeviceToHost), "read hits");
        } catch (...) {
            hipFree(d_r);
            if (d_h) hipFree(d_h);
            throw;
        }
        hipFree(d_r);
        hipFree(d_h);
    });
}

int rt_build_photon_grid_from(rt_renderer *r, const rt_photon *photons, uint32_t n_slots) {
    return api_call([&] {
        if (!r || (n_slots && !photons)) throw ApiError(RT_ERR_INVALID, "null argument");
        if (!r->cfg.enable_photon_map) throw ApiError(RT_ERR_INVALID, "photon map disabled in the config");
        if (n_slots > r->photon_slots) throw ApiError(RT_ERR_CAPACITY, "more slots than the photon buffer holds");
        hip_check(hipStreamSynchronize(r->stream), "sync");
        hip_check(hipMemset(r->d_photons, 0, sizeof(rt_photon) * r->photon_slots), "clear photons");
        if (n_slots)
            hip_check(hipMemcpy(r->d_photons, photons, sizeof(rt_photon) * n_slots, hipMemcpyHostToDevice), "upload");
        r->grid_build(n_slots, nullptr);
    });
}

int rt_photon_grid_build_noalloc(const rt_photon *d_photons, uint32_t n_slots, const rt_grid_params *gp,
                                 uint32_t *d_grid_meta, uint32_t meta_cells, rt_photon *d_sorted,
                                 float *d_gather_rec, uint32_t *d_n_sorted, void *stream) {
    return api_call([&] {
        if (!gp || !d_n_sorted || (n_slots && (!d_photons || !d_sorted)) || (meta_cells && !d_grid_meta))
            throw ApiError(RT_ERR_INVALID, "null argument");
        if (n_slots > RT_GRID_NOALLOC_MAX_SLOTS)
            throw ApiError(RT_ERR_CAPACITY, "n_slots " + std::to_string(n_slots) +
                                                " exceeds the allocation-free workspace of " +
                                                std::to_string(RT_GRID_NOALLOC_MAX_SLOTS) + " slots");
        if (meta_cells < gp->num_cells) throw ApiError(RT_ERR_INVALID, "meta_cells below gp->num_cells");
        rtk::GridBuildLaunch g;
        g.photons = d_photons;
        g.n_slots = n_slots;
        g.gp = *gp;
        g.meta_cells = meta_cells;
        g.grid_meta = d_grid_meta;
        g.sorted = d_sorted;
        g.gather_rec = d_gather_rec;
        g.temp = nullptr;  // the library's fixed workspace
        g.temp_bytes = 0;
        g.n_valid_out = d_n_sorted;
        hip_check(rtk::launch_grid_build_noalloc(g, static_cast<hipStream_t>(stream)), "allocation-free grid build");
    });
}

int rt_get_pass_stats(rt_renderer *r, rt_pass_stats *out) {
    return api_call([&] {
        if (!r || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        hip_check(hipStreamSynchronize(r->stream), "sync");
        r->collect();
        *out = r->stats;
    });
}

int rt_reset_pass_stats(rt_renderer *r) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        hip_check(hipStreamSynchronize(r->stream), "sync");
        r->collect();
        std::memset(&r->stats, 0, sizeof(r->stats));
    });
}

int rt_count_gather_traffic(rt_renderer *r, rt_trace_counters *out) {
    return api_call([&] {
        if (!r || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        if (r->cfg.enable_photon_map && !r->photons_built)
            throw ApiError(RT_ERR_INVALID, "photon map enabled but not built");
        hip_check(hipMemsetAsync(r->d_counters, 0, sizeof(unsigned long long) * kCounterWords, r->stream),
                  "clear counters");
        r->ensure_path_queues();
        const uint32_t saved = r->sample_counter;
        r->sample_counter = saved + 1u;  // the frame the next rt_render_frame would trace
        rtk::GatherLaunch g = r->gather_launch(1.0f, r->d_ping, r->d_ping, r->d_ping, (uint32_t)rt_renderer::kSets);
        r->sample_counter = saved;
        g.counters = r->d_counters;
        hip_check(rtk::launch_gather(g, r->stream), "gather (counting)");
        unsigned long long c[kCounterWords];
        read_dev(r, c, r->d_counters, sizeof(c));
        out->ray_queries = c[0];
        out->nodes_visited = c[1];
        out->tris_tested = c[2];
        out->photon_reads = c[3];
        for (int k = 0; k < RT_COUNT_PHASES; k++) {
            out->lane_steps[k] = c[4 + k];
            out->wave_steps[k] = c[4 + RT_COUNT_PHASES + k];
            out->wave_cycles[k] = c[4 + 2 * RT_COUNT_PHASES + k];
        }
    });
}

// ───────────────────────────── many-view batch ─────────────────────────────

static void batch_build_tiles(rt_renderer &r, uint32_t tw, uint32_t th) {
    std::vector<rtk::BatchViewRecord> recs(r.views.size());
    struct T {
        uint32_t view, tile, cost;
    };
    std::vector<T> tiles;
    for (size_t v = 0; v < r.views.size(); v++) {
        const rt_view &vw = r.views[v];
        rtk::BatchViewRecord &rec = recs[v];
        std::memset(&rec, 0, sizeof(rec));
        rec.cam = vw.camera;
        rec.width = vw.width;
        rec.height = vw.height;
        rec.spp = vw.spp;
        rec.frame = vw.frame;
        rec.pixel_offset = r.view_offsets[v];
        rec.tiles_x = (vw.width + tw - 1) / tw;
        const uint32_t ty = (vw.height + th - 1) / th;
        for (uint32_t k = 0; k < rec.tiles_x * ty; k++) tiles.push_back({(uint32_t)v, k, std::max(vw.spp, 1u)});
    }
    // longest-processing-time first: the queue hands out the costliest tiles first
    std::stable_sort(tiles.begin(), tiles.end(), [](const T &a, const T &b) { return a.cost > b.cost; });
    std::vector<uint32_t> flat(tiles.size() * 2);
    for (size_t i = 0; i < tiles.size(); i++) {
        flat[2 * i] = tiles[i].view;
        flat[2 * i + 1] = tiles[i].tile;
    }
    hip_check(hipStreamSynchronize(r.stream), "sync");
    if (recs.size() > r.views_cap) {
        if (r.d_views) hip_check(hipFree(r.d_views), "free");
        r.d_views = dev_alloc<rtk::BatchViewRecord>(recs.size(), "batch views");
        r.views_cap = recs.size();
    }
    if (flat.size() > r.tiles_cap) {
        if (r.d_tiles) hip_check(hipFree(r.d_tiles), "free");
        r.d_tiles = dev_alloc<uint32_t>(flat.size(), "batch tiles");
        r.tiles_cap = flat.size();
    }
    hip_check(hipMemcpy(r.d_views, recs.data(), recs.size() * sizeof(rtk::BatchViewRecord), hipMemcpyHostToDevice),
              "upload views");
    hip_check(hipMemcpy(r.d_tiles, flat.data(), flat.size() * 4, hipMemcpyHostToDevice), "upload tiles");
    r.n_tiles = (uint32_t)tiles.size();
    r.tiles_w = tw;
    r.tiles_h = th;
}

static void batch_validate(const rt_batch_tuning &t) {
    if (t.tile_w == 0 || t.tile_h == 0 || (t.tile_w * t.tile_h) % 64u != 0 || t.tile_w * t.tile_h > 4096u)
        throw ApiError(RT_ERR_INVALID, "tile_w * tile_h must be a multiple of 64 (at most 4096)");
    if (t.waves_per_cu == 0 || t.waves_per_cu > 20)
        throw ApiError(RT_ERR_INVALID, "waves_per_cu must be in 1..20 (8 KB of LDS stack per wavefront)");
    if (t.tiles_per_fetch == 0 || t.tiles_per_fetch > 64) throw ApiError(RT_ERR_INVALID, "tiles_per_fetch in 1..64");
    if (t.lanes > 1u) throw ApiError(RT_ERR_INVALID, "lanes must be 0 or 1");
}

static void batch_launch(rt_renderer &r, const rt_batch_tuning &t) {
    if (r.views.empty()) throw ApiError(RT_ERR_INVALID, "no views set (rt_batch_set_views)");
    if (r.cfg.enable_photon_map && !r.photons_built)
        throw ApiError(RT_ERR_INVALID, "photon map enabled but rt_rebuild_photons has not run");
    batch_validate(t);
    if (t.tile_w != r.tiles_w || t.tile_h != r.tiles_h) batch_build_tiles(r, t.tile_w, t.tile_h);
    rtk::GatherLaunch g = r.gather_launch(1.0f, r.d_bframe, r.d_bmom, r.d_bgb, (uint32_t)rt_renderer::kSets);
    g.px_state = nullptr;  // the batch kernel is one persistent pass
    rtk::BatchLaunch b;
    b.views = r.d_views;
    b.tiles = r.d_tiles;
    b.n_tiles = r.n_tiles;
    b.tile_w = t.tile_w;
    b.tile_h = t.tile_h;
    b.tiles_per_fetch = t.tiles_per_fetch;
    b.n_workgroups = (uint32_t)r.n_cu * t.waves_per_cu;
    b.counter = r.d_counter;
    b.lanes = t.lanes;
    if (t.lanes && (uint64_t)r.n_tiles * t.tile_w * t.tile_h >= (1ull << 32))
        throw ApiError(RT_ERR_CAPACITY, "batch has 2^32 or more pixel slots");
    hipEvent_t ev;
    r.begin(RT_PASS_BATCH, ev);
    hip_check(rtk::launch_gather_batch(g, b, r.stream), "gather batch");
    r.end(RT_PASS_BATCH, ev);
}

extern "C" int rt_batch_set_views(rt_renderer *r, const rt_view *views, uint32_t n) {
    return api_call([&] {
        if (!r || (n && !views)) throw ApiError(RT_ERR_INVALID, "null argument");
        uint64_t px = 0, samples = 0;
        std::vector<uint64_t> offs(n);
        for (uint32_t v = 0; v < n; v++) {
            if (views[v].width == 0 || views[v].height == 0 || views[v].width > 16384 || views[v].height > 16384)
                throw ApiError(RT_ERR_INVALID, "view resolution must be in 1..16384");
            offs[v] = px;
            px += (uint64_t)views[v].width * views[v].height;
            samples += (uint64_t)views[v].width * views[v].height * std::max(views[v].spp, 1u);
        }
        if (px > (1ull << 30)) throw ApiError(RT_ERR_CAPACITY, "batch exceeds 2^30 pixels");
        hip_check(hipStreamSynchronize(r->stream), "sync");
        if (px > r->batch_cap_px) {
            for (float **b : {&r->d_bframe, &r->d_bmom, &r->d_bgb}) {
                if (*b) hip_check(hipFree(*b), "free");
                *b = dev_alloc<float>(px * 4, "batch outputs");
            }
            r->batch_cap_px = px;
        }
        if (!r->d_counter) r->d_counter = dev_alloc<uint32_t>(1, "batch counter");
        if (r->n_cu == 0) {
            hip_check(hipDeviceGetAttribute(&r->n_cu, hipDeviceAttributeMultiprocessorCount, r->device), "CU count");
        }
        r->views.assign(views, views + n);
        r->view_offsets = offs;
        r->batch_px = px;
        r->batch_samples = samples;
        r->tiles_w = r->tiles_h = 0;  // rebuild the tile table at the next launch
    });
}

extern "C" int rt_render_batch(rt_renderer *r, const rt_batch_tuning *tuning) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        if (tuning) {
            batch_validate(*tuning);
            r->tuning = *tuning;
        }
        batch_launch(*r, r->tuning);
    });
}

extern "C" int rt_batch_autotune(rt_renderer *r, uint32_t repeats, rt_batch_tuning *best, double *best_ms) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        repeats = std::max(repeats, 1u);
        const uint32_t shapes[][2] = {{8, 8}, {16, 4}, {16, 16}, {32, 8}};
        const uint32_t waves[] = {4, 8, 12, 16};
        // (lanes, tiles per fetch): pixel slots per lane, or 1 or 4 tiles per wavefront
        const uint32_t forms[][2] = {{1, 1}, {0, 1}, {0, 4}};
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        rt_batch_tuning winner = r->tuning;
        double winner_ms = 1e30;
        try {
            for (auto &sh : shapes) {
                for (uint32_t w : waves) {
                    for (auto &f : forms) {
                        const rt_batch_tuning t{sh[0], sh[1], w, f[1], f[0]};
                        batch_launch(*r, t);  // warm-up (also builds the tile table)
                        hip_check(hipEventRecord(e0, r->stream), "event");
                        for (uint32_t k = 0; k < repeats; k++) batch_launch(*r, t);
                        hip_check(hipEventRecord(e1, r->stream), "event");
                        hip_check(hipEventSynchronize(e1), "sync");
                        float ms = 0.f;
                        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
                        const double per = ms / repeats;
                        if (per < winner_ms) {
                            winner_ms = per;
                            winner = t;
                        }
                    }
                }
            }
        } catch (...) {
            hipEventDestroy(e0);
            hipEventDestroy(e1);
            throw;
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        r->collect();
        r->tuning = winner;
        if (best) *best = winner;
        if (best_ms) *best_ms = winner_ms;
    });
}

extern "C" int rt_batch_get_tuning(const rt_renderer *r, rt_batch_tuning *out) {
    return api_call([&] {
        if (!r || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        *out = r->tuning;
    });
}

extern "C" int rt_batch_totals(const rt_renderer *r, uint64_t *pixels, uint64_t *samples) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        if (pixels) *pixels = r->batch_px;
        if (samples) *samples = r->batch_samples;
    });
}

extern "C" int rt_read_batch_view(rt_renderer *r, uint32_t view, float *frame, float *moments, float *gbuffer,
                                  size_t n_floats) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        if (view >= r->views.size()) throw ApiError(RT_ERR_INVALID, "view index out of range");
        const size_t npx = (size_t)r->views[view].width * r->views[view].height;
        if (n_floats < npx * 4) throw ApiError(RT_ERR_INVALID, "buffer smaller than width*height*4 floats");
        const size_t off = r->view_offsets[view] * 4;
        if (frame) read_dev(r, frame, r->d_bframe + off, npx * 16);
        if (moments) read_dev(r, moments, r->d_bmom + off, npx * 16);
        if (gbuffer) read_dev(r, gbuffer, r->d_bgb + off, npx * 16);
    });
}

int rt_hbm_copy_probe(int device, size_t bytes, uint32_t iters, double *gbps, double *ms_per_copy) {
    return api_call([&] {
        if (!gbps || bytes < 16 || iters == 0) throw ApiError(RT_ERR_INVALID, "bad argument");
        select_device(device);
        void *src = nullptr, *dst = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            if (src) hipFree(src);
            if (dst) hipFree(dst);
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
            if (s) hipStreamDestroy(s);
        };
        try {
            bytes &= ~(size_t)15;
            hip_check(hipMalloc(&src, bytes), "probe alloc");
            hip_check(hipMalloc(&dst, bytes), "probe alloc");
            hip_check(hipMemset(src, 1, bytes), "probe fill");
            hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "probe stream");
            hip_check(hipEventCreate(&e0), "probe event");
            hip_check(hipEventCreate(&e1), "probe event");
            hip_check(rtk::launch_hbm_copy(src, dst, bytes, s), "probe warmup");
            hip_check(hipEventRecord(e0, s), "probe event");
            for (uint32_t i = 0; i < iters; i++) hip_check(rtk::launch_hbm_copy(src, dst, bytes, s), "probe copy");
            hip_check(hipEventRecord(e1, s), "probe event");
            hip_check(hipEventSynchronize(e1), "probe sync");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "probe elapsed");
            const double per = ms / iters;
            *gbps = 2.0 * (double)bytes / (per * 1e-3) / 1e9;
            if (ms_per_copy) *ms_per_copy = per;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

int rt_accum64_reset(rt_renderer *r) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        if (!r->d_acc64) r->d_acc64 = dev_alloc<double>(r->n_px() * 4, "fp64 sum");
        hip_check(hipMemsetAsync(r->d_acc64, 0, r->n_px() * 32, r->stream), "clear fp64 sum");
        r->acc64_frames = 0;
    });
}

int rt_accum64_add_accumulation(rt_renderer *r) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        if (!r->d_acc64) throw ApiError(RT_ERR_INVALID, "call rt_accum64_reset first");
        hipEvent_t ev;
        r->begin(RT_PASS_ACCUM64, ev);
        hip_check(rtk::launch_accum64_add(r->d_accum, r->d_acc64, r->n_px(), r->stream), "accum64");
        r->end(RT_PASS_ACCUM64, ev);
        r->acc64_frames++;
    });
}

int rt_accum64_relmse(rt_renderer *r, const void *d_image, float eps, double *relmse) {
    return api_call([&] {
        if (!r || !relmse) throw ApiError(RT_ERR_INVALID, "null argument");
        if (!r->d_acc64 || r->acc64_frames == 0) throw ApiError(RT_ERR_INVALID, "the FP64 sum holds no frames");
        if (!(eps > 0.0f)) throw ApiError(RT_ERR_INVALID, "eps must be positive");
        const size_t nd = rtk::relmse_scratch_doubles();
        if (!r->d_relmse) r->d_relmse = dev_alloc<double>(nd, "relmse scratch");
        const void *img = d_image ? d_image : r->d_accum;
        hip_check(rtk::launch_relmse(img, r->d_acc64, r->acc64_frames, eps, r->n_px(), r->d_relmse, r->stream),
                  "relmse");
        read_dev(r, relmse, r->d_relmse + (nd - 1), sizeof(double));
    });
}

int rt_accum64_relmse_display(rt_renderer *r, float eps, double *relmse) {
    if (!r) return rt_accum64_relmse(r, nullptr, eps, relmse);
    return rt_accum64_relmse(r, r->unet ? r->d_display : nullptr, eps, relmse);
}

int rt_renderer_set_denoiser(rt_renderer *r, rt_unet *u) {
    return api_call([&] {
        if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
        hip_check(hipStreamSynchronize(r->stream), "sync");
        if (u && !r->d_display) {
            r->d_display = dev_alloc<float>(r->n_px() * 4, "display");
            hip_check(hipMemset(r->d_display, 0, r->n_px() * 16), "memset");
        }
        r->unet = u;
        r->ensure_mom_hist();
    });
}

int rt_read_display(rt_renderer *r, float *rgba, size_t n) {
    return read_px4(r, rgba, n, r ? (r->unet ? r->d_display : r->d_accum) : nullptr);
}

int rt_read_albedo(rt_renderer *r, float *rgba, size_t n) {
    return api_call([&] {
        if (!r || !rgba) throw ApiError(RT_ERR_INVALID, "null argument");
        r->ensure_albedo();
        if (n < r->n_px() * 4) throw ApiError(RT_ERR_INVALID, "buffer smaller than width*height*4");
        read_dev(r, rgba, r->d_albedo, r->n_px() * 16);
    });
}

int rt_accum64_read_mean(rt_renderer *r, double *rgba, size_t n, uint64_t *n_frames) {
    return api_call([&] {
        if (!r || !rgba) throw ApiError(RT_ERR_INVALID, "null argument");
        if (!r->d_acc64) throw ApiError(RT_ERR_INVALID, "call rt_accum64_reset first");
        if (n < r->n_px() * 4) throw ApiError(RT_ERR_INVALID, "buffer smaller than width*height*4 doubles");
        read_dev(r, rgba, r->d_acc64, r->n_px() * 32);
        const double k = r->acc64_frames ? 1.0 / (double)r->acc64_frames : 0.0;
        for (size_t i = 0; i < r->n_px() * 4; i++) rgba[i] *= k;
        if (n_frames) *n_frames = r->acc64_frames;
    });
}

}  // extern "C"

namespace rtu {

FrameViews frame_views(rt_renderer *r) {
    if (!r) throw ApiError(RT_ERR_INVALID, "null renderer");
    r->ensure_albedo();
    hip_check(hipStreamSynchronize(r->stream), "sync");
    return r->frame_views();
}

}  // namespace rtu
