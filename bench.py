#!/usr/bin/env python3
"""Benchmark of the rendering path on one MI355X (gfx950).

Workload (default): the full 1920x1080 frame on the ~1M-triangle synthetic
scene (600k-triangle value-noise terrain + 200 instanced 2k-triangle meshes,
seed 1), 1 sample per pixel per frame, path depth 8, photon map of 1M photons
traced once before timing (the reference's trace-once rebuild), a-trous on
(2 passes), running-mean accumulation. One step = one frame: gather + 2 a-trous
passes + composite/tonemap. The timed frames are issued with rt_render_frames,
as the headless renderer renders a frame sequence (eight frames per camera-pass
launch); the rate of frames issued one rt_render_frame (record_frame) call at a
time is reported beside it. Inputs (scene, BVH, photons) are resident in HBM
before the timed region.

Prints one JSON line. `value` is Msamples/s (camera paths per second). The
roofline block is for the dominant kernel (gather): algorithmic bytes per
launch from the traversal counters of one extra, untimed launch, divided by the
kernel's mean duration measured with HIP events on the renderer's stream.
The cpu_baseline is the CPU oracle (a C++ restatement of the same WGSL, OpenMP)
rendering whole frames of the same workload (gather only); its first frame is
checked bit for bit against the GPU's moments and G-buffer of that frame.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s @1080p; ms to relMSE≤0.01: batched+learned vs plain à-trous"
HBM_PEAK_GBS = 8000.0
CUS, CLOCK_GHZ = 256, 2.4  # MI355X compute units and peak engine clock
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16

# Algorithmic bytes per unit of work of the traversal layouts: a frame's bytes are
# 32 per child box tested + 48 per triangle tested + 32 per ray query (ray + hit)
# + 36 per photon candidate + 48 of outputs per pixel.
NODE_BOX_BYTES = 32    # one child box: min/max (binary record 64 B / 2, 4-wide record 128 B / 4)
TRI_BYTES = 48         # slot triangle: v0, v1-v0, v2-v0 (3 x float4)
RAY_BYTES = 32         # 24 B ray + 8 B hit
PHOTON_BYTES = 36      # per radius-gather candidate: position + power (16), power + incident (16), incident.z (4)
PIXEL_OUT_BYTES = 48   # frame + moments + gbuffer float4 writes
PRESENT_REPEATS = 5
# the learned denoiser's parameters, trained on the 2000-scene training workload by
# `python -m rustracer_amd.denoiser --scenes 2000 --steps 2000 --out <this file>`
CONFIG4_WEIGHTS = "rustracer_amd/weights/unet_config4.f32"
DENOISE_PX_BYTES = 64  # per pass: color in + gbuffer + moments + out, 16 B each
COMPOSITE_TONEMAP_PX_BYTES = 52  # frame 16 + accumulation 16 read + 16 written + rgba8 4
GRID_SLOT_BYTES = 32   # count (position + power.r, 16) + scatter key read (16) per slot
GRID_STORED_BYTES = 128  # stored photon: 64 B record read + 64 B written
GRID_CELL_BYTES = 8    # [start, count] per cell
# The reference's own layout, from the oracle's counters of the same frame: a
# 48-byte BVHNode per node visited, a 128-byte TriangleGpu per triangle tested,
# a 64-byte PhotonGpu per photon candidate (bvh.rs:79-88, scene.rs:74-94,
# lib.rs:86-97).
REF_NODE_BYTES, REF_TRI_BYTES, REF_PHOTON_BYTES = 48, 128, 64


TRAVERSALS = {0: "reference BVH2 nodes", 1: "binary child-box records, reference order",
              2: "binary child-box records, nearest first", 3: "4-wide records, reference order",
              4: "4-wide records, reference order, watertight triangle test"}


PROFILE_FRAMES = 8     # frames of the per-pass kernel times (one rt_render_frames launch)
CHUNK = 8              # frames per camera-pass launch of rt_render_frames (16 output sets)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--photons", type=int, default=1_000_000)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--reference-traversal", action="store_true")
    ap.add_argument("--nearest-first", action="store_true",
                    help="child-box traversal nearest child first (not bit-exact by construction)")
    ap.add_argument("--watertight", action="store_true",
                    help="watertight ray/triangle test instead of Moller-Trumbore (not bit-exact by construction)")
    ap.add_argument("--binary-records", action="store_true",
                    help="binary child-box records instead of the 4-wide records (both exact)")
    ap.add_argument("--path-queues", action="store_true",
                    help="camera pass as path-queue passes instead of one kernel per frame (same results)")
    ap.add_argument("--tile-launch", action="store_true",
                    help="camera pass as one workgroup per 8x8 tile instead of resident wavefronts whose lanes "
                         "take new pixels (same results)")
    ap.add_argument("--no-photons", action="store_true")
    ap.add_argument("--record-frame", action="store_true",
                    help="time frames issued one rt_render_frame call at a time for the headline value")
    ap.add_argument("--no-record-frame", action="store_true",
                    help="issue no frame one call at a time (every camera-pass launch then renders eight "
                         "frames when --steps is a multiple of eight; for profiler runs)")
    ap.add_argument("--no-network", action="store_true", help="skip the learned denoiser's timing")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-view batch measurement")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 720p photon-map workload and the network training throughput")
    ap.add_argument("--convergence-train", action="store_true",
                    help="train a small network inside the run (64 scenes, 800 steps) for the convergence study "
                         "instead of using the shipped weights trained on the 2000-scene workload")
    ap.add_argument("--no-convergence", action="store_true",
                    help="skip the ms-to-relMSE study (about two minutes: a 16384-frame FP64 reference, a dataset "
                         "and a training run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding pixels, views or photon indices across GPUs is designed, not built)", file=sys.stderr)
        return 2

    import numpy as np

    import rustracer_amd as rt
    from rustracer_amd import _ffi

    t0 = time.perf_counter()
    scene = rt.Scene.builtin("fullframe", 1)
    cam = rt.Scene.builtin_camera("fullframe", 1)
    t_scene = time.perf_counter() - t0
    counts = scene.counts()
    log(f"scene fullframe: {counts['triangles']} triangles ({t_scene:.1f}s)")
    flags = ((_ffi.FLAG_REFERENCE_TRAVERSAL if args.reference_traversal else 0) |
             (_ffi.FLAG_NEAREST_FIRST if args.nearest_first else 0) |
             (_ffi.FLAG_BINARY_RECORDS if args.binary_records else 0) |
             (_ffi.FLAG_WATERTIGHT if args.watertight else 0) |
             (_ffi.FLAG_PATH_QUEUES if args.path_queues else 0) |
             (_ffi.FLAG_TILE_LAUNCH if args.tile_launch else 0))
    photons_on = not args.no_photons
    cfg = rt.RenderConfig(width=args.width, height=args.height, spp=1, max_depth=args.depth,
                          enable_photon_map=photons_on, photon_count=args.photons, enable_denoise=True,
                          accumulate_frames=1 << 20, flags=flags)
    t0 = time.perf_counter()
    r = rt.Renderer(cfg, scene, 0)
    t_create = time.perf_counter() - t0
    tinfo = r.traversal_info()
    log(f"renderer created (flatten + BVH build + upload) in {t_create:.1f}s; traversal {tinfo}")
    r.update_camera(cam)
    base_cfg = r.config

    def set_profile(on: bool) -> None:
        c = r.config
        c.flags = base_cfg.flags | (_ffi.FLAG_PROFILE if on else 0)
        r.config = c

    set_profile(True)
    t0 = time.perf_counter()
    if photons_on:
        r.rebuild_photons()
    r.synchronize()
    t_photons = time.perf_counter() - t0
    photon_stats = r.pass_stats()
    r.reset_pass_stats()
    set_profile(False)
    log(f"photon map rebuilt in {t_photons*1e3:.1f} ms")

    issued = 0  # frames rendered so far: the output-set rotation is at issued mod 16

    def timed_frames(n: int, one_call: bool) -> float:
        nonlocal issued
        issued += n
        r.synchronize()
        t0 = time.perf_counter()
        if one_call:
            r.render_frames(n)
        else:
            for _ in range(n):
                r.record_frame()
        r.synchronize()
        return time.perf_counter() - t0

    if args.record_frame and args.no_record_frame:
        log("--record-frame and --no-record-frame exclude each other")
        return 2
    if not args.no_record_frame:
        timed_frames(args.warmup, False)
    timed_frames(args.warmup, True)
    log(f"warmup {args.warmup} frames done; timing {args.steps} frames")

    # timed region: frames back to back (consecutive camera passes overlap on
    # two streams; rt_render_frames also renders eight frames per launch)
    elapsed = timed_frames(args.steps, not args.record_frame)
    elapsed_other = None if args.no_record_frame else timed_frames(args.steps, args.record_frame)
    log(f"{args.steps} frames in {elapsed * 1e3:.1f} ms" +
        ("" if elapsed_other is None else f" ({elapsed_other * 1e3:.1f} ms the other way)"))

    # per-pass kernel times, each launch on its own (no overlap), for the
    # rooflines: frames issued as in the timed region, starting at a multiple of
    # eight in the output-set rotation (so rt_render_frames makes one launch)
    if not args.no_record_frame:
        for _ in range((-issued) % CHUNK):
            r.record_frame()
            issued += 1
    r.synchronize()
    set_profile(True)
    r.reset_pass_stats()
    if args.record_frame:
        for _ in range(PROFILE_FRAMES):
            r.record_frame()
            r.synchronize()
    else:
        r.render_frames(PROFILE_FRAMES)
        r.synchronize()
    issued += PROFILE_FRAMES
    stats = r.pass_stats()
    w, h = args.width, args.height
    # the present pass (fullscreen.wgsl) of the last output into a 4K surface and
    # into one of the render size: mean kernel time over a few launches
    present_ms = {}
    for sw, sh in ((3840, 2160), (w, h)):
        r.reset_pass_stats()
        for _ in range(PRESENT_REPEATS):
            r.read_present(sw, sh)
        ms_, n_ = r.pass_stats()["present"]
        present_ms[(sw, sh)] = ms_ / max(n_, 1)
    set_profile(False)
    samples = w * h * args.steps
    msps = samples / elapsed / 1e6
    ms_per_step = elapsed / args.steps * 1e3

    # traversal traffic of the next frame (one untimed counting launch)
    cnt = r.count_gather_traffic()
    gather_ms, gather_n = stats["gather"]
    gather_avg_ms = gather_ms / max(gather_n, 1)
    frames_per_launch = PROFILE_FRAMES // max(gather_n, 1)
    # algorithmic bytes of one frame (counted on the next frame) x the frames of one launch
    frame_bytes = (cnt["nodes_visited"] * NODE_BOX_BYTES + cnt["tris_tested"] * TRI_BYTES +
                   cnt["ray_queries"] * RAY_BYTES + cnt["photon_reads"] * PHOTON_BYTES + w * h * PIXEL_OUT_BYTES)
    alg_bytes = frame_bytes * frames_per_launch
    achieved = alg_bytes / (gather_avg_ms * 1e-3) / 1e9
    mrays = cnt["ray_queries"] * args.steps / elapsed / 1e6

    result = {
        "metric": METRIC,
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "fullframe-1080p: 1M-triangle terrain+instances, 1 spp/frame, depth "
                               f"{args.depth}, {args.photons} photons, a-trous x2",
                   "width": w, "height": h, "spp": 1, "depth": args.depth, "photons": args.photons,
                   "triangles": counts["triangles"],
                   "traversal": TRAVERSALS[tinfo["mode"]], "bvh_depth": tinfo["bvh_depth"],
                   "camera_pass": ("path queues (primary pass + one compacted pass per depth)" if args.path_queues else
                                   "one workgroup per tile" if args.tile_launch else
                                   "persistent lanes taking new pixels"),
                   "photon_map": photons_on,
                   "frames_issued": ("one rt_render_frame call per frame" if args.record_frame else
                                     "rt_render_frames (eight frames per camera-pass launch)")},
        "mrays_per_s": round(mrays, 2),
        "pass_ms_per_launch": {k: round(v[0] / max(v[1], 1), 4) for k, v in stats.items() if v[1]},
        "gather_frames_per_launch": frames_per_launch,
        "gather_ms_per_frame": round(gather_ms / PROFILE_FRAMES, 4),
        "pass_ms_note": (f"kernel times of {PROFILE_FRAMES} frames issued as in the timed region, each launch on "
                         "its own (HIP events on each pass's stream); a-trous and composite launch once per "
                         "frame (a-trous twice)"),
        "photon_rebuild_ms": round(t_photons * 1e3, 2),
        "photon_pass_ms": {k: round(v[0], 3) for k, v in photon_stats.items() if v[1]},
        "setup_s": {"scene": round(t_scene, 2), "create_bvh_upload": round(t_create, 2)},
        "gather_traffic_per_frame": cnt,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "gather", "kernel_avg_ms": round(gather_avg_ms, 4),
                     "frames_per_launch": frames_per_launch,
                     "algorithmic_bytes_per_launch": int(alg_bytes)},
        "cpu_baseline": None,
    }

    # streaming passes: algorithmic bytes per launch / mean launch time
    gp = r.grid_params()
    pass_bytes = {"denoise": w * h * DENOISE_PX_BYTES,
                  "composite_tonemap": w * h * COMPOSITE_TONEMAP_PX_BYTES}
    result["pass_rooflines"] = {}
    for name, nbytes in pass_bytes.items():
        ms_, n_ = stats.get(name, (0.0, 0))
        if n_:
            gbs = nbytes / (ms_ / n_ * 1e-3) / 1e9
            result["pass_rooflines"][name] = {"bound": "hbm", "achieved": round(gbs, 1), "unit": "GB/s",
                                              "frac": round(gbs / HBM_PEAK_GBS, 4),
                                              "bytes_per_launch": int(nbytes)}
    for (sw, sh), ms_ in present_ms.items():
        nbytes = w * h * 4 + sw * sh * 4  # rgba8 output read once, bgra8 surface written once
        gbs = nbytes / (ms_ * 1e-3) / 1e9
        result["pass_rooflines"][f"present_{sw}x{sh}"] = {"bound": "hbm", "achieved": round(gbs, 1), "unit": "GB/s",
                                                           "frac": round(gbs / HBM_PEAK_GBS, 4),
                                                           "bytes_per_launch": int(nbytes), "ms": round(ms_, 4)}
    if photons_on and photon_stats.get("grid_build", (0, 0))[1]:
        srt = r.read_sorted_photons()
        n_slots = min(max(args.photons * 6 // 256, 1) * 256, gp.photon_slots)
        nbytes = n_slots * GRID_SLOT_BYTES + len(srt) * GRID_STORED_BYTES + gp.num_cells * GRID_CELL_BYTES
        ms_, n_ = photon_stats["grid_build"]
        gbs = nbytes / (ms_ / n_ * 1e-3) / 1e9
        result["pass_rooflines"]["grid_build"] = {"bound": "hbm", "achieved": round(gbs, 1), "unit": "GB/s",
                                                  "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": int(nbytes),
                                                  "slots": n_slots, "stored": len(srt)}

    # HBM traffic of the gather dispatch, measured with rocprofv3 PMC counters on this
    # workload (FETCH_SIZE / WRITE_SIZE, scaled against the copy probe) and kept
    # with the profiles; reported only when the configuration is the one measured
    traffic_file = os.path.join(ROOT, "profiles", "gather_fullframe_1080p_traffic.txt")
    default_run = (w, h, args.depth, args.photons, photons_on) == (1920, 1080, 8, 1_000_000, True)
    if default_run and os.path.exists(traffic_file):
        t = json.load(open(traffic_file))
        form = "path_queues" if args.path_queues else "tile_launch" if args.tile_launch else "persistent"
        if (t.get("traversal_mode") == tinfo["mode"] and t.get("camera_pass") == form and
                t.get("frames_per_launch", 1) == frames_per_launch):
            result["roofline"]["traffic"] = t["hbm_bytes_per_launch"]
            result["roofline"]["traffic_source"] = t["source"]
            result["roofline"]["note"] = ("the algorithmic bytes are served from L2/MALL (HBM traffic is the "
                                          "'traffic' figure); VALU issue bounds this kernel: "
                                          "profiles/gather_fullframe_1080p_valu_counters.md")

    # VALU issue of the gather launch, from rocprofv3 counters of this workload
    # (SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU, SQ_THREAD_CYCLES_VALU) kept with the
    # profiles: the bound the camera pass actually meets
    valu_file = os.path.join(ROOT, "profiles", "gather_fullframe_1080p_valu.json")
    if default_run and os.path.exists(valu_file):
        v = json.load(open(valu_file))
        form = "path_queues" if args.path_queues else "tile_launch" if args.tile_launch else "persistent"
        if (v.get("traversal_mode") == tinfo["mode"] and v.get("camera_pass") == form and
                v.get("frames_per_launch") == frames_per_launch):
            # one wave64 VALU instruction per cycle per CU (4 SIMDs, 4 cycles each)
            issue_ms = v["SQ_INSTS_VALU"] / CUS / (CLOCK_GHZ * 1e6)
            lanes = v["SQ_THREAD_CYCLES_VALU"] / v["SQ_ACTIVE_INST_VALU"]
            result["valu_roofline"] = {
                "bound": "valu", "kernel": "gather", "valu_instructions_per_launch": v["SQ_INSTS_VALU"],
                "issue_ms_at_peak": round(issue_ms, 4), "kernel_avg_ms": round(gather_avg_ms, 4),
                "issue_frac": round(issue_ms / gather_avg_ms, 4), "active_lanes_per_valu": round(lanes, 2),
                "lane_frac": round(lanes / 64.0, 4), "source": v["source"],
                "note": f"issue at one wave64 VALU instruction per cycle per CU, {CUS} CUs at {CLOCK_GHZ} GHz; the "
                        "HBM roofline above counts algorithmic bytes that L2/MALL serve"}
            result["roofline"]["note"] = ("bound by VALU issue, not HBM: see valu_roofline; the algorithmic bytes "
                                          "are served from L2/MALL ('traffic' is the measured HBM traffic)")

    # where the camera pass's time goes: shares of the wavefronts' cycles per
    # phase, from the shader clock read at the phase boundaries in the counting
    # launch (bounce = closest-hit traversal + shading + shadow traversal +
    # photon search; the loop adds the pixel hand-out and sample ends)
    cyc = cnt.get("wave_cycles", {})
    if cyc.get("camera_loop"):
        tot = cyc["camera_loop"]
        inner = cyc["traversal"] + cyc["shadow_traversal"] + cyc["photon_search"]
        result["gather_cycle_share"] = {
            "closest_traversal": round(cyc["traversal"] / tot, 4),
            "shadow_traversal": round(cyc["shadow_traversal"] / tot, 4),
            "photon_search": round(cyc["photon_search"] / tot, 4),
            "rest_of_bounce": round((cyc["shading"] - inner) / tot, 4),
            "refill_and_sample_end": round((tot - cyc["shading"]) / tot, 4)}

    if elapsed_other is not None:
        key = "msamples_per_s_render_frames" if args.record_frame else "msamples_per_s_record_frame"
        result[key] = round(w * h * args.steps / elapsed_other / 1e6, 3)

    copy_gbs, copy_ms = rt.hbm_copy_probe(0, 1 << 30, 20)
    result["roofline"]["measured_copy_gbs"] = round(copy_gbs, 1)
    result["roofline"]["frac_of_measured_copy"] = round(achieved / copy_gbs, 4)
    log(f"HBM copy probe: {copy_gbs:.0f} GB/s (1 GiB copy in {copy_ms:.3f} ms)")
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(rt, r, scene, cam, cfg, args)
        oc = result["cpu_baseline"].pop("oracle_counters_first_frame")
        # the same frame's algorithmic bytes in the reference's layout, from the
        # oracle's counters of the reference traversal
        ref_bytes = (oc["nodes_visited"] * REF_NODE_BYTES + oc["tris_tested"] * REF_TRI_BYTES +
                     oc["ray_queries"] * RAY_BYTES + oc["photon_reads"] * REF_PHOTON_BYTES + w * h * PIXEL_OUT_BYTES)
        result["reference_layout_bytes"] = {
            "frame_bytes_reference_layout": int(ref_bytes), "frame_bytes_this_layout": int(frame_bytes),
            "oracle_counters": oc,
            "same_ray_queries": oc["ray_queries"] == cnt["ray_queries"],
            "note": "the frame the counting launch traced, counted again by the oracle's reference traversal "
                    "(48 B nodes, 128 B triangles, 64 B photon records)"}
    if not args.no_network and w % 2 == 0 and h % 2 == 0:
        result["network_roofline"] = network_roofline(r, w, h)
    r.close()
    if not args.no_batch:
        result["batch_multiview"] = batch_multiview(args)
    if not args.no_extra:
        result["photon_plain_720p"] = photon_plain_720p()
        if not args.no_network:
            result["network_training"] = network_training()
    if not args.no_convergence:
        result["convergence"] = convergence(args.convergence_train)
    print(json.dumps(result), flush=True)
    return 0


def batch_multiview(args) -> dict:
    """64 seeded views of the ~109k-triangle multiview scene (log-uniform
    128x128 .. 1920x1080, spp in {1, 2, 4, 8, 16}, depth 8) in one persistent
    launch: the autotuner picks tile shape, waves per CU and tiles per fetch,
    then the launch is timed."""
    import rustracer_amd as rt
    from rustracer_amd.workloads import multiview_views

    scene = rt.Scene.builtin("multiview", 1)
    views = multiview_views(64, seed=1)
    cfg = rt.RenderConfig(width=8, height=8, spp=1, max_depth=8, enable_denoise=False)
    with rt.Renderer(cfg, scene, 0) as r:
        r.set_views(views)
        px, smp = r.batch_totals()
        r.render_batch()
        r.synchronize()
        best, tuned_ms = r.batch_autotune(repeats=1)
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            r.render_batch()
        r.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
    log(f"batch: 64 views, {px / 1e6:.1f} Mpx, {smp / 1e6:.1f} Msamples in {ms:.1f} ms (tuning {best})")
    return {"workload": "64 views of the multiview scene (109k triangles), log-uniform 128^2..1920x1080, "
                        "spp in {1,2,4,8,16}, depth 8, one persistent launch",
            "pixels": px, "samples": smp, "ms_per_launch": round(ms, 3),
            "msamples_per_s": round(smp / (ms * 1e-3) / 1e6, 2), "tuning": best,
            "autotune_best_ms": round(tuned_ms, 3)}


def photon_plain_720p() -> dict:
    """The photon-mapped plain scene at 1280x720 (displaced GGX-metal icosphere, glass
    sphere, Lambert floor, area + point light): 1M photons, grid cell = radius = 0.05,
    4 spp x 32 accumulated frames, depth 6, a-trous x2."""
    import rustracer_amd as rt

    scene = rt.Scene.builtin("photon_plain", 1)
    cam = rt.Scene.builtin_camera("photon_plain", 1)
    cfg = rt.RenderConfig(width=1280, height=720, spp=4, max_depth=6, enable_photon_map=True,
                          photon_count=1_000_000, grid_cell_size=0.05, enable_denoise=True, accumulate_frames=32)
    with rt.Renderer(cfg, scene, 0) as r:
        r.update_camera(cam)
        t0 = time.perf_counter()
        r.rebuild_photons()
        r.synchronize()
        t_ph = time.perf_counter() - t0
        r.render_frames(2)
        r.synchronize()
        r.reset_accumulation()
        t0 = time.perf_counter()
        r.render_frames(32)
        r.synchronize()
        el = time.perf_counter() - t0
        gp = r.grid_params()
    smp = 1280 * 720 * 4 * 32
    log(f"photon_plain 720p: 32 frames x 4 spp in {el * 1e3:.1f} ms; photon rebuild {t_ph * 1e3:.1f} ms")
    return {"workload": "photon-mapped plain scene 1280x720, 1M photons, cell = radius = 0.05, 4 spp x 32 frames, "
                        "depth 6, a-trous x2", "msamples_per_s": round(smp / el / 1e6, 2),
            "ms_per_frame": round(el / 32 * 1e3, 3), "photon_rebuild_ms": round(t_ph * 1e3, 2),
            "grid_cells": gp.num_cells, "triangles": scene.counts()["triangles"]}


def network_training() -> dict:
    """Training throughput of the learned denoiser at the training workload's
    batch shape: Adam steps on batches of 32 crops of 256x256 from a small
    renderer-made dataset. FLOPs per step = forward + weight gradients +
    data gradients (every layer but the first) of the 3x3 layers, plus the
    output layer's."""
    from rustracer_amd.denoiser import UNet, build_dataset

    net = UNet(0, 1)
    build_dataset(net, n_scenes=4, size=256, target_samples=256, photons=50_000, first_seed=77)
    net.train(steps=3, batch=32, crop=256, lr=1e-3, seed=1)
    steps = 20
    t0 = time.perf_counter()
    net.train(steps=steps, batch=32, crop=256, lr=1e-3, seed=2)
    el = time.perf_counter() - t0
    net.close()
    flops = network_training_flops(256, 256) * 32
    tf = flops * steps / el / 1e12
    log(f"network training: {el / steps * 1e3:.2f} ms/step (batch 32 x 256^2), {tf:.1f} TFLOP/s")
    return {"bound": "mfma", "ms_per_step": round(el / steps * 1e3, 3), "batch": "32 x 256x256",
            "flops_per_step": int(flops), "achieved": round(tf, 1), "peak": MFMA_BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4),
            "note": "wall time per Adam step incl. the crop gather and the host's per-step crop draw"}


def network_flops(w: int, h: int) -> float:
    """Multiply-add FLOPs of one U-Net forward pass at w x h (padded to multiples of 16)."""
    from rustracer_amd.denoiser import network_flops as nf
    return nf(w, h)


def network_training_flops(w: int, h: int) -> float:
    """One crop's training FLOPs: 3x the forward's, less the first layer's
    data gradient (the features need none)."""
    from rustracer_amd.denoiser import LAYER_SHAPES
    first = 2.0 * w * h * 9 * 12 * LAYER_SHAPES[0][1]
    return 3.0 * network_flops(w, h) - first


def network_roofline(r, w: int, h: int) -> dict:
    """The learned denoiser's pass on this frame size (weights untrained; the
    cost does not depend on their values): mean time of the network pass from
    HIP events, against MI355X's dense BF16 MFMA peak."""
    from rustracer_amd import _ffi
    from rustracer_amd.denoiser import UNet
    net = UNet(0, 1)
    r.set_denoiser(net)
    c = r.config
    c.flags = c.flags | _ffi.FLAG_PROFILE
    r.config = c
    r.record_frame()  # albedo pass and first use
    r.synchronize()
    r.reset_pass_stats()
    for _ in range(PROFILE_FRAMES):
        r.record_frame()
        r.synchronize()
    ms, n = r.pass_stats()["network"]
    r.set_denoiser(None)
    net.close()
    avg = ms / max(n, 1)
    flops = network_flops(w, h)
    tf = flops / (avg * 1e-3) / 1e12
    log(f"network pass {avg:.3f} ms at {w}x{h}: {tf:.1f} TFLOP/s")
    return {"bound": "mfma", "achieved": round(tf, 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4), "ms_per_frame": round(avg, 4),
            "flops_per_frame": int(flops), "note": "whole pass: features, 18 BF16 MFMA 3x3 convolutions "
                                                   "(five levels, 32-128 channels), pools, output layer, tonemap"}


def convergence(train_in_run: bool = False) -> dict:
    """ms to relMSE <= 0.01 (learned U-Net vs a-trous) on this workload, measured
    now by `python -m rustracer_amd.convergence` in a child process: a
    16384-frame FP64 reference, then each method's frames until the displayed
    image reaches relMSE 0.01, re-run back to back and timed. The network is the
    shipped one trained on the 2000-scene training workload (CONFIG4_WEIGHTS),
    or with train_in_run a small one trained inside the run (64 scenes, 800 steps)."""
    import subprocess
    cmd = [sys.executable, "-m", "rustracer_amd.convergence"]
    if not train_in_run and os.path.exists(os.path.join(ROOT, CONFIG4_WEIGHTS)):
        cmd += ["--weights", CONFIG4_WEIGHTS]
    out = subprocess.run(cmd, cwd=ROOT, check=True, capture_output=True, text=True).stdout
    d = json.loads(out.strip().splitlines()[-1])
    methods = {m["method"]: m for m in d["methods"]}
    res = {"metric": "ms to relMSE<=0.01 (eps 1e-2) vs FP64 16384-frame reference", "source": "measured in this run",
           "methods": {k: {"frames": v.get("frames"), "ms": v.get("ms")} for k, v in methods.items()}}
    for k in ("speedup_vs_atrous", "network", "training"):
        if k in d:
            res["speedup_learned_vs_atrous" if k == "speedup_vs_atrous" else k] = d[k]
    return res


def cpu_baseline(rt, r, scene, cam, cfg, args):
    """The oracle on host cores over whole frames, from the frame the GPU just
    rendered on, until about --cpu-seconds of CPU time."""
    import numpy as np

    from tests.helpers import image_params
    from tests.oracle_ffi import OracleScene

    # every core this process may run on, unless OMP_NUM_THREADS caps the share
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, max(1, int(os.environ["OMP_NUM_THREADS"])))
    gp, meta, srt = r.grid_params(), r.read_grid_meta(), r.read_sorted_photons()
    # the gather's own outputs (the frame buffer itself holds the a-trous result)
    gpu_mom, gpu_gb = r.read_moments(), r.read_gbuffer()
    c = r.config
    frame_idx = r.sample_count()
    o = OracleScene.from_scene(scene)
    u = rt.camera_to_uniform(cam, c.width, c.height)
    frames, elapsed, exact, first_stats = 0, 0.0, None, None
    while (elapsed < args.cpu_seconds and frames < 16) or frames < 2:
        ip = image_params(c, frame_idx + frames, alpha=1.0)
        t0 = time.perf_counter()
        _, mom, gb, st = o.gather(u, ip, max_depth=c.max_depth, photon_map=(gp, meta, srt), threads=threads)
        elapsed += time.perf_counter() - t0
        if frames == 0:  # the frame the GPU rendered last
            exact = bool(np.array_equal(mom, gpu_mom) and np.array_equal(gb, gpu_gb))
        if frames == 1:  # the frame rt_count_gather_traffic traced
            first_stats = {k: int(getattr(st, k)) for k in ("ray_queries", "nodes_visited", "tris_tested",
                                                             "photon_reads")}
        frames += 1
    msps = frames * c.width * c.height / elapsed / 1e6
    log(f"cpu baseline: {frames} full frames on {threads} threads in {elapsed:.1f}s, "
        f"moments+gbuffer bit-exact vs GPU: {exact}")
    return {"value": round(msps, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full {c.width}x{c.height} frames x 1 spp (gather only, same photon map)",
            "sample_bit_exact_vs_gpu": exact, "cores_note": "the cores this process may use (sched_getaffinity), "
            "capped by OMP_NUM_THREADS", "oracle_counters_first_frame": first_stats}


if __name__ == "__main__":
    sys.exit(main())
