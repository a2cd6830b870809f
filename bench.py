"""Benchmarks of the LBVH build + traversal core on one GPU.

Default workload "all": the "frame" line below, with the "batched" and
"serve" workloads' key figures added under "batched" and "serve" (the
metric's batched-vs-plain and p50/p99 parts).

Workload "frame" (one step = one frame): 262,144 triangles
(centroids U[0,1]^3, edges N(0, 0.004^2)), rigidly rotated 1 degree per frame
about the axis (1, 2, 3) through (0.5, 0.5, 0.5); the LBVH is rebuilt every
frame (bounds, Morton, onesweep sort, emit, refit), then a 1920x1080 frame
traces one primary ray per pixel and one shadow ray per lit-side hit to a
point light.  value = (primary + shadow rays) / summed device time of the
render kernel (HIP events on the library's stream); build_ms is the device
time of one rebuild.  Scene data is resident in device memory before the
timed region; nothing is copied per frame.

Other workloads (--workload), each printing one JSON line of the same shape:
  sphere   2,097,152-triangle noise-displaced sphere, 8 orbit cameras, 4 lights
           (one step = the 8 frames); value = Mrays/s, plus HBM GB/s.
  batched  64 object-mix scenes (1,024..200,000 triangles, log-uniform), images
           of 256^2/512^2/1024^2, diffuse GI at 4 spp x 4 bounces; one step =
           rebuild all 64 + render all 64, batched (one build, one persistent
           launch; also as a replayed graph) against per-scene builds and
           launches; images compared bit for bit.
  whitted  the frame's 262,144-triangle soup (static) with a seeded mix of matte,
           mirror, glass and half-mirror materials, 1920x1080 Whitted frames
           of 3 bounces (primary, closest-hit continuation and shadow rays, the
           reference's ray-tracing colour stage); value = Mrays/s.
  bvh      the frame's soup (static) traced through the SAH BVH (host build, as
           the reference builds it) and through the LBVH: direct frames at
           1920x1080; value = Mrays/s on the BVH, with the LBVH's rate, both
           structures' SAH costs and build times beside it.
  build    LBVH rebuilds of triangle soups of 262,144 to 33,554,432 triangles:
           per-stage device times, Mtris/s, and the onesweep sort's achieved
           GB/s (4 passes x (8 B read + 8 B written) per key) against 8 TB/s.
  serve    Poisson 200 requests/s for 60 s over a pool of 500 scenes
           (1k..500k triangles), 512^2 or 1920x1080, direct or 2-bounce GI at
           1 spp with progressive accumulation, 200 GB device budget; p50/p99
           latency and aggregate Mrays/s against one request per launch (the
           server with a 1 ms batch window and centre-first tiles, and without
           the window).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "raytracer-gpupro4_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "Mrays/s + LBVH build ms @262k tris; batched vs plain; p50/p99 latency"
CPU_SHARE = 16  # host threads a one-GPU box grants a process

NTRI, SIGMA, SEED, DEPTH = 262144, 0.004, 2, 20
W, H = 1920, 1080
CAM_POS, CAM_TGT, FOV = (0.5, 0.5, -1.5), (0.5, 0.5, 0.5), 60.0
LIGHTS = [(0.5, 2.5, -1.0)]
AXIS, CENTER = (1.0, 2.0, 3.0), (0.5, 0.5, 0.5)
# algorithmic bytes per traced ray and per test (ray 48 B in, hit 24 B out,
# 32 B node box per box test, 4 B id + 12 B indices + 36 B vertices per
# triangle test)
RAY_BYTES, BOX_BYTES, TRI_BYTES = 48 + 24, 32, 52
# launch shape of the headline frame (block, LDS stack entries, blocks per CU,
# traversal form, quad tiles): held fixed so that the profile of the same
# command times the same kernel.  rt_autotune_direct (on the first, static
# frame) picks this shape or (64, 20, ., 2, 64), 1-2 % apart on the rotating
# frames, so its choice flipped from run to run; --autotune searches again.
FRAME_TUNING = (128, 20, 8, 2, 0)
PMC_SUMMARY = os.path.join(ROOT, "profiles", "render_direct_soup262k_1080p_pmc.md")
SPHERE_PMC_SUMMARY = os.path.join(ROOT, "profiles", "render_direct_quad_sphere2m_1080p_pmc.md")


_T0 = time.perf_counter()


def log(msg):
    print(f"[bench {time.perf_counter() - _T0:6.1f} s] {msg}", file=sys.stderr, flush=True)


def depth_for(nt):
    """Heap depth for nt triangles: two levels below a balanced binary tree."""
    return int(min(24, max(8, math.ceil(math.log2(max(nt, 2))) + 2)))


def view_camera(rtcore, xyz, view, w, h, dist=1.6):
    """Camera on a circle around the scene's box, view k at 30 + 90k degrees."""
    lo, hi = xyz.min(0), xyz.max(0)
    c = 0.5 * (lo + hi)
    r = 0.5 * float(np.linalg.norm(hi - lo))
    a = math.radians(30.0 + 90.0 * view)
    pos = c + dist * r * np.array([math.cos(a), 0.45, math.sin(a)], np.float32)
    return rtcore.camera_look_at(tuple(pos), tuple(c), (0, 1, 0), 60.0, w / h)


def result(metric, value, unit, steps, warmup, wall, config, dtype="f32", **extra):
    out = {"metric": metric, "value": round(value, 2), "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": round(wall * 1e3 / max(steps, 1), 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    return out


def traffic_from_profile(path=None, kernel=None):
    """HBM bytes per launch of the frame kernel measured by rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE, scaled by their copy-kernel calibration),
    read from a committed summary (the line for `kernel` where the summary
    covers several); None when it is absent."""
    key = "traffic_bytes_per_launch" + (f" {kernel}" if kernel else "") + ":"
    try:
        for line in open(path or PMC_SUMMARY):
            if line.startswith(key):
                return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


# ------------------------------------------------------------------ frame
def direct_kernel(tuning):
    """The kernel that renders direct frames under this tuning: with quads
    (tuning[4] > 0) the heaviest tiles and the rest share one launch of
    k_render_direct_quad."""
    return "k_render_direct_quad" if tuning[4] > 0 else "k_render_direct"


def kernel_name(tuning):
    """The frame kernel's instance as rocprofv3 names it, e.g.
    k_render_direct<false, false, 128, 20, 2> (block, LDS stack, form)."""
    if tuning[4] > 0:
        return f"k_render_direct_quad<{tuning[0]}, {tuning[1]}, {tuning[3]}>"
    return f"k_render_direct<false, false, {tuning[0]}, {tuning[1]}, {tuning[3]}>"


def frame(ctx, rtcore, s, cam, f, flags=0, stats=False):
    ctx.scene_transform(s, rtcore.rotation_about(AXIS, float(f), CENTER))
    ctx.build(s)
    return ctx.render_direct(s, cam, W, H, LIGHTS, flags=flags, want_stats=stats)


def cpu_baseline(rtcore, xyz, idx, cam, budget_s=12.0):
    """The oracle doing the same per-frame work on the host: build of the
    rotated scene on 1 thread (the reference builds on one CPU thread) and
    closest-hit primary + shadow traversal of every pixel on up to CPU_SHARE
    threads, frame after frame until budget_s seconds have been spent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    affinity = len(os.sched_getaffinity(0))
    cores = min(CPU_SHARE, affinity)
    f32 = np.float32
    rays0 = oracle.raygen_pinhole(cam.as_array(), W, H)
    build_s = trace_s = 0.0
    nrays = frames = 0
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < budget_s and frames < 64:
        m = rtcore.rotation_about(AXIS, float(frames), CENTER)
        d = (xyz - m[9:12]).astype(f32)
        R = m[:9].reshape(3, 3)
        rot = np.stack([((((R[r, 0] * d[:, 0]).astype(f32) + (R[r, 1] * d[:, 1]).astype(f32)).astype(f32)
                          + (R[r, 2] * d[:, 2]).astype(f32)).astype(f32) + m[9 + r]).astype(f32)
                        for r in range(3)], 1)
        t0 = time.perf_counter()
        ref = oracle.build_scene(rot, idx, DEPTH, presorted=False)
        build_s += time.perf_counter() - t0
        if frames == 0:
            ref0, rot0 = ref, rot
        t0 = time.perf_counter()
        hits = oracle.trace(ref["nodes"], ref["ids"], rot, idx, rays0, nthreads=cores)
        _, occ = oracle.render_direct(ref["nodes"], ref["ids"], rot, idx, rays0, hits, LIGHTS, nthreads=cores)
        trace_s += time.perf_counter() - t0
        nrays += len(rays0) + int((occ != 2).sum())
        frames += 1
    # the same traversal on one thread, over every 16th pixel of frame 0
    sub = rays0[::16].copy()
    t0 = time.perf_counter()
    hits1 = oracle.trace(ref0["nodes"], ref0["ids"], rot0, idx, sub, nthreads=1)
    _, occ1 = oracle.render_direct(ref0["nodes"], ref0["ids"], rot0, idx, sub, hits1, LIGHTS, nthreads=1)
    one_s = time.perf_counter() - t0
    one_rays = len(sub) + int((occ1 != 2).sum())
    return {"value": round(nrays / trace_s / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"{frames} full frames of the same workload ({W}x{H} primary + shadow rays, "
                      f"closest hit, shadow rays from the hit leaf): {nrays} rays in {trace_s:.1f} s "
                      f"on {cores} threads; oracle rebuild of {NTRI} tris per frame on 1 thread",
            "value_1thread": round(one_rays / one_s / 1e6, 3),
            "sample_1thread": f"every 16th pixel of frame 0: {one_rays} rays in {one_s:.1f} s",
            "build_ms_1thread": round(build_s * 1e3 / frames, 2), "cpu_model": cpu_model(),
            "affinity_cores": affinity,
            "threads_note": f"traversal threads = min({CPU_SHARE}, affinity mask): a one-GPU box grants a "
                            f"process {CPU_SHARE} host threads whatever the mask shows"}


def cpu_model():
    """The host CPU as /proc/cpuinfo names it: model name, family/model/
    stepping, vendor and the logical CPUs online."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = line.split(":", 1)
                info.setdefault(k.strip(), v.strip())
    except OSError:
        return "unknown"
    return (f"{info.get('model name', 'unknown')} ({info.get('vendor_id', '?')} family {info.get('cpu family', '?')} "
            f"model {info.get('model', '?')} stepping {info.get('stepping', '?')}, {os.cpu_count()} logical CPUs, "
            f"{info.get('cpu MHz', '?')} MHz)")


def build_end_to_end(ctx, s, n=5):
    """Mean device time of n rebuilds, each timed by one event before it
    and one after (a profiling window), in ms."""
    ctx.profile_begin()
    for _ in range(n):
        ctx.build(s)
    prof = ctx.profile_end()
    return prof["build_ms"]["total"] / prof["builds"]


def build_stages(ctx, s, n=5):
    """Median per-stage device times of n instrumented rebuilds (an event at
    every stage boundary; each stalls the stream about 5 us, so their total
    is above the uninstrumented build_ms, which a profiling window times with
    one event before the build and one after)."""
    times = [ctx.build(s, timed=True) for _ in range(n)]
    out = {k.replace("_ms", ""): round(float(np.median([t[k] for t in times])), 4) for k in times[0]}
    out["note"] = "instrumented rebuilds: one event per stage boundary"
    return out


def device_sync(ctx):
    """Waits for the library's stream, then for the whole device."""
    import torch

    ctx.sync()
    torch.cuda.synchronize()


def workload_frame(args, rtcore):
    xyz, idx = rtcore.gen_triangle_soup(NTRI, SIGMA, SEED)
    cam = rtcore.camera_look_at(CAM_POS, CAM_TGT, (0, 1, 0), FOV, W / H)
    ctx = rtcore.Context(0)
    s = ctx.add_scene(xyz, idx, DEPTH)
    ctx.build(s)
    if args.autotune:
        tuned = ctx.autotune_direct(s, cam, W, H, LIGHTS)  # launch shape for this frame, before timing
    else:
        ctx.set_tuning(*FRAME_TUNING)
        tuned = ctx.tuning()
    for f in range(args.warmup):
        frame(ctx, rtcore, s, cam, f)
    device_sync(ctx)

    ctx.profile_begin()
    t0 = time.perf_counter()
    for f in range(args.steps):
        frame(ctx, rtcore, s, cam, f)
    device_sync(ctx)
    wall = time.perf_counter() - t0
    prof = ctx.profile_end()

    # counting pass over the same frames (outside the timed region): box and
    # triangle tests per frame for the algorithmic byte count
    boxes = tris = rays_counted = 0
    for f in range(args.steps):
        st = frame(ctx, rtcore, s, cam, f, flags=rtcore.FLAG_COUNT, stats=True)
        boxes += st["box_tests"]
        tris += st["tri_tests"]
        rays_counted += st["primary_rays"] + st["shadow_rays"]

    rays = prof["primary_rays"] + prof["shadow_rays"]
    assert rays == rays_counted, (rays, rays_counted)
    trace_ms = prof["trace_ms"]
    mrays = rays / (trace_ms * 1e-3) / 1e6
    alg_bytes = RAY_BYTES * rays + BOX_BYTES * boxes + TRI_BYTES * tris
    per_launch_bytes = alg_bytes / prof["launches"]
    per_launch_ms = trace_ms / prof["launches"]
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    build_total = prof["build_ms"]["total"] / prof["builds"]
    stages = build_stages(ctx, s)
    ctx.close()
    traffic = traffic_from_profile(PMC_SUMMARY, kernel_name(tuned))
    out = result(METRIC, mrays, "Mrays/s", args.steps, args.warmup, wall,
                 {"workload": f"{NTRI}-triangle soup rebuilt every frame (1 deg/frame rotation), "
                              f"{W}x{H} primary + shadow rays, 1 point light",
                  "tris": NTRI, "depth": DEPTH, "width": W, "height": H, "lights": len(LIGHTS)},
                 build_ms=round(build_total, 4),
                 build_ms_stages=stages,
                 trace_ms_per_frame=round(per_launch_ms, 4),
                 rays_per_frame=rays / args.steps,
                 shadow_rays_per_frame=prof["shadow_rays"] / args.steps,
                 frame_mrays_s_wall=round(rays / wall / 1e6, 2),
                 tests_per_ray={"box": boxes / rays, "tri": tris / rays},
                 tuning=list(tuned), tuning_source="autotuned" if args.autotune else "fixed (FRAME_TUNING)",
                 roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                           "traffic": traffic, "kernel": kernel_name(tuned),
                           "bytes_per_launch": per_launch_bytes,
                           "note": "achieved = algorithmic bytes (per-ray box and triangle tests x their record "
                                   "sizes) / launch time; most of them are served by L2 and the Infinity Cache, "
                                   "traffic is what reaches HBM"},
                 cpu_baseline=None)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(rtcore, xyz, idx, cam)
    return out


# ------------------------------------------------------------------ sphere
def workload_sphere(args, rtcore):
    xyz, idx = rtcore.gen_noisy_sphere(1025, 1024, 0.05, 3)
    nt = len(idx)
    depth = depth_for(nt)
    lights = [(3.0, 3.0, 3.0), (-3.0, 3.0, 3.0), (3.0, 3.0, -3.0), (-3.0, -3.0, -3.0)]
    cams = []
    for k in range(8):
        a = 2.0 * math.pi * k / 8
        cams.append(rtcore.camera_look_at((2.5 * math.cos(a), 0.0, 2.5 * math.sin(a)), (0, 0, 0), (0, 1, 0),
                                          60.0, W / H))
    ctx = rtcore.Context(0)
    s = ctx.add_scene(xyz, idx, depth)
    ctx.build(s)
    bt = {"total_ms": build_end_to_end(ctx, s)}
    tuned = ctx.autotune_direct(s, cams[0], W, H, lights)
    for _ in range(args.warmup):
        for c in cams:
            ctx.render_direct(s, c, W, H, lights, want_stats=False)
    ctx.sync()
    ctx.profile_begin()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for c in cams:
            ctx.render_direct(s, c, W, H, lights, want_stats=False)
    ctx.sync()
    wall = time.perf_counter() - t0
    prof = ctx.profile_end()
    boxes = tris = 0
    for c in cams:
        st = ctx.render_direct(s, c, W, H, lights, flags=rtcore.FLAG_COUNT)
        boxes += st["box_tests"]
        tris += st["tri_tests"]
    ctx.close()
    rays = prof["primary_rays"] + prof["shadow_rays"]
    per_step_rays = rays / args.steps
    mrays = rays / (prof["trace_ms"] * 1e-3) / 1e6
    alg = RAY_BYTES * per_step_rays + BOX_BYTES * boxes + TRI_BYTES * tris  # per step (8 frames)
    per_launch_ms = prof["trace_ms"] / prof["launches"]
    achieved = (alg / 8) / (per_launch_ms * 1e-3) / 1e9
    return result(METRIC, mrays, "Mrays/s", args.steps, args.warmup, wall,
                  {"workload": f"{nt}-triangle noise-displaced sphere (static), 8 orbit cameras "
                               f"(radius 2.5) at {W}x{H}, primary + shadow rays to 4 point lights",
                   "tris": nt, "depth": depth, "cameras": 8, "lights": 4},
                  build_ms=round(bt["total_ms"], 4), trace_ms_per_frame=round(per_launch_ms, 4),
                  tuning=list(tuned),
                  rays_per_frame=per_step_rays / 8,
                  tests_per_ray={"box": boxes / per_step_rays, "tri": tris / per_step_rays},
                  roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                            "traffic": traffic_from_profile(SPHERE_PMC_SUMMARY, direct_kernel(tuned)),
                            "kernel": direct_kernel(tuned), "bytes_per_launch": alg / 8},
                  cpu_baseline=None)


# ------------------------------------------------------------------ whitted
WHITTED_BOUNCES = 3


def whitted_materials(rtcore, nt, seed=17):
    """Matte, mirror, glass (index 1.5), half mirror and the default
    MaterialData, assigned to triangles uniformly at random."""
    m = np.zeros(5, rtcore.MATERIAL_DTYPE)
    m[0] = ((0.8, 0.6, 0.4, 1.0), (0.1, 0.1, 0.1, 1.0), (0.2, 0.2, 0.2, 1.0), 0.0, 1.0, 0, 0)
    m[1] = ((0.3, 0.3, 0.3, 0.2), (0.0, 0.0, 0.0, 1.0), (0.9, 0.9, 0.95, 1.0), 0.5, 1.0, 1, 0)
    m[2] = ((0.95, 0.97, 0.9, 0.6), (0.02, 0.02, 0.02, 1.5), (0.8, 0.8, 0.8, 0.7), 0.9, 0.0, 2, 0)
    m[3] = ((0.2, 0.7, 0.3, 0.5), (0.05, 0.1, 0.05, 1.2), (0.6, 0.6, 0.6, 1.0), 0.3, 0.5, 3, 0)
    m[4] = ((0.5, 0.5, 0.5, 0.8), (0.0, 0.0, 0.0, 1.0), (0.8, 0.8, 0.8, 1.0), 0.0, 0.0, 4, 0)
    ids = np.random.default_rng(seed).integers(0, 5, nt).astype(np.uint32)
    return m, ids


def workload_whitted(args, rtcore):
    xyz, idx = rtcore.gen_triangle_soup(NTRI, SIGMA, SEED)
    mats, mat_ids = whitted_materials(rtcore, len(idx))
    cam = rtcore.camera_look_at(CAM_POS, CAM_TGT, (0, 1, 0), FOV, W / H)
    kw = dict(bounces=WHITTED_BOUNCES, switches=rtcore.WHITTED_SHADOWS | rtcore.WHITTED_GLOSS,
              env=(0.4, 0.5, 0.6), seed=5)
    ctx = rtcore.Context(0)
    s = ctx.add_scene(xyz, idx, DEPTH)
    ctx.build(s)
    ctx.scene_shading(s, None, mat_ids, mats)
    tuned, kw["switches"] = ctx.autotune_whitted(s, cam, W, H, LIGHTS[0], **kw)
    for f in range(args.warmup):
        ctx.render_whitted(s, cam, W, H, LIGHTS[0], frame=f, want_stats=False, **kw)
    device_sync(ctx)
    ctx.profile_begin()
    t0 = time.perf_counter()
    for f in range(args.steps):
        ctx.render_whitted(s, cam, W, H, LIGHTS[0], frame=f, want_stats=False, **kw)
    device_sync(ctx)
    wall = time.perf_counter() - t0
    prof = ctx.profile_end()
    boxes = tris = counted = 0
    for f in range(args.steps):
        st = ctx.render_whitted(s, cam, W, H, LIGHTS[0], frame=f, flags=rtcore.FLAG_COUNT, **kw)
        boxes += st["box_tests"]
        tris += st["tri_tests"]
        counted += st["primary_rays"] + st["shadow_rays"] + st["bounce_rays"]
    # the other device forms over the same frames (same images and rays)
    forms = {"wavefront": 0, "one-pass": rtcore.WHITTED_ONE_PASS, "split": rtcore.WHITTED_SPLIT,
             "split-quads": rtcore.WHITTED_SPLIT | rtcore.WHITTED_QUADS}
    base_sw = kw["switches"] & ~(rtcore.WHITTED_ONE_PASS | rtcore.WHITTED_SPLIT | rtcore.WHITTED_QUADS)
    form = next(k for k, bit in forms.items() if kw["switches"] == base_sw | bit)
    others = []
    for name, bit in forms.items():
        if name == form:
            continue
        alt = dict(kw, switches=base_sw | bit)
        ctx.profile_begin()
        for f in range(args.steps):
            ctx.render_whitted(s, cam, W, H, LIGHTS[0], frame=f, want_stats=False, **alt)
        alt_prof = ctx.profile_end()
        others.append({"form": name, "trace_ms_per_frame": round(alt_prof["trace_ms"] / alt_prof["launches"], 4),
                       "mrays_s": round((alt_prof["primary_rays"] + alt_prof["shadow_rays"] + alt_prof["bounce_rays"])
                                        / (alt_prof["trace_ms"] * 1e-3) / 1e6, 2)})
    rays = prof["primary_rays"] + prof["shadow_rays"] + prof["bounce_rays"]
    assert rays == counted, (rays, counted)
    mrays = rays / (prof["trace_ms"] * 1e-3) / 1e6
    per_launch_ms = prof["trace_ms"] / prof["launches"]
    alg = (RAY_BYTES * rays + BOX_BYTES * boxes + TRI_BYTES * tris) / prof["launches"]
    achieved = alg / (per_launch_ms * 1e-3) / 1e9
    ctx.close()
    kernels = {"wavefront": "k_whitted_first + k_whitted_next per bounce", "one-pass": "k_render_whitted",
               "split": "k_whitted_first + k_whitted_rest",
               "split-quads": "k_whitted_first + k_whitted_rest_quad"}
    out = result(METRIC, mrays, "Mrays/s", args.steps, args.warmup, wall,
                 {"workload": f"{NTRI}-triangle soup (static), 5 materials (matte, mirror, glass 1.5, "
                              f"half mirror, default), {W}x{H} Whitted frames, {WHITTED_BOUNCES} bounces, "
                              f"shadow rays to 1 point light",
                  "tris": NTRI, "depth": DEPTH, "width": W, "height": H, "bounces": WHITTED_BOUNCES},
                 form=form, trace_ms_per_frame=round(per_launch_ms, 4), other_forms=others,
                 rays_per_frame=rays / args.steps,
                 shadow_rays_per_frame=prof["shadow_rays"] / args.steps,
                 bounce_rays_per_frame=prof["bounce_rays"] / args.steps,
                 tests_per_ray={"box": boxes / rays, "tri": tris / rays}, tuning=list(tuned),
                 roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                           "kernel": kernels[form], "bytes_per_launch": alg},
                 cpu_baseline=None)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        cores = min(CPU_SHARE, len(os.sched_getaffinity(0)))
        ref = oracle.build_scene(xyz, idx, DEPTH)
        sw, sh = W // 4, H // 4  # the same view at a quarter of the resolution per axis
        scam = rtcore.camera_look_at(CAM_POS, CAM_TGT, (0, 1, 0), FOV, sw / sh)
        t0 = time.perf_counter()
        _, _, _, n = oracle.render_whitted(ref["nodes"], ref["ids"], xyz, idx, scam.as_array(), sw, sh, LIGHTS[0],
                                           mat_ids=mat_ids, materials=mats, nthreads=cores, **kw)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n / dt / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
                               "sample": f"one {sw}x{sh} Whitted frame of the same scene, view and materials: "
                                         f"{n} rays in {dt:.1f} s on {cores} threads"}
    return out


# ------------------------------------------------------------------ build
BUILD_SIZES = (262144, 2097152, 8388608, 33554432)


def workload_build(args, rtcore):
    sizes = [args.tris] if args.tris else list(BUILD_SIZES)
    ctx = rtcore.Context(0)
    rows = []
    for nt in sizes:
        xyz, idx = rtcore.gen_triangle_soup(nt, SIGMA * (262144 / nt) ** (1 / 3), SEED)
        depth = depth_for(nt)
        s = ctx.add_scene(xyz, idx, depth)
        del xyz, idx
        for _ in range(max(args.warmup, 1)):
            ctx.build(s)
        ctx.sync()
        # end to end (an event before and after each build), then the stages
        # of instrumented rebuilds
        ctx.profile_begin()
        for _ in range(args.steps):
            ctx.build(s)
        prof = ctx.profile_end()
        total = prof["build_ms"]["total"] / prof["builds"]
        times = [ctx.build(s, timed=True) for _ in range(args.steps)]
        med = {k: float(np.median([t[k] for t in times])) for k in times[0]}
        sort_bytes = 4 * 16 * nt
        rows.append({"tris": nt, "depth": depth, "build_ms": round(total, 4),
                     "stages_ms": {k.replace("_ms", ""): round(v, 4) for k, v in med.items()},
                     "mtris_s": round(nt / (total * 1e-3) / 1e6, 1),
                     "sort_gb_s": round(sort_bytes / (med["sort_ms"] * 1e-3) / 1e9, 1),
                     "sort_frac": round(sort_bytes / (med["sort_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
        log(f"build {nt}: {rows[-1]}")
        ctx.scene_destroy(s)
    ctx.close()
    first, last = rows[0], rows[-1]
    return result(METRIC, first["build_ms"], "ms", args.steps, args.warmup, first["build_ms"] * args.steps * 1e-3,
                  {"workload": "LBVH rebuild of seeded triangle soups (edges scaled to keep density), "
                               + ", ".join(str(r["tris"]) for r in rows) + " triangles",
                   "tris": [r["tris"] for r in rows]},
                  higher_is_better=False, builds=rows,
                  roofline={"bound": "hbm", "achieved": last["sort_gb_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": last["sort_frac"], "traffic": None,
                            "kernel": "sort passes (k_sort_count + k_sort_scan + k_sort_scatter at this size)",
                            "bytes_per_launch": 16 * last["tris"], "tris": last["tris"]},
                  cpu_baseline=None)


# ------------------------------------------------------------------ bvh
def sah_cost(nodes, heap):
    """Sum over reachable nodes of area / root area x (0.125 inner, primitive
    count leaf): the same measure for the LBVH heap (root 1, children 2n and
    2n + 1, leaf bit 4, count << 3) and the flattened BVH (root 0)."""
    def area(n):
        d = (n["mx"] - n["mn"]).astype(np.float64)
        return 2.0 * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2])

    root = 1 if heap else 0
    ra = area(nodes[root])
    cost, stack = 0.0, [root]
    while stack:
        i = stack.pop()
        n = nodes[i]
        pc = int(n["prim_count"])
        leaf = (pc & 4) != 0 if heap else pc > 0
        if leaf:
            cost += area(n) / ra * ((pc >> 3) if heap else pc)
        else:
            cost += area(n) / ra * 0.125
            stack += [2 * i, 2 * i + 1] if heap else [i + 1, int(n["prim_pos"])]
    return cost


def workload_bvh(args, rtcore):
    xyz, idx = rtcore.gen_triangle_soup(NTRI, SIGMA, SEED)
    cam = rtcore.camera_look_at(CAM_POS, CAM_TGT, (0, 1, 0), FOV, W / H)
    ctx = rtcore.Context(0)
    s = ctx.add_scene(xyz, idx, DEPTH)
    ctx.build(s)
    lb = {"total_ms": build_end_to_end(ctx, s)}
    heap = ctx.read(s)[0]
    lbvh_cost = sah_cost(heap, True)
    rates = {}
    for name in ("lbvh", "bvh"):
        if name == "bvh":
            info = ctx.bvh_build(s)
            nodes, _ = ctx.bvh_read(s, info["nodes"])
            bvh_cost = sah_cost(nodes, False)
        tuned = ctx.autotune_direct(s, cam, W, H, LIGHTS)
        for _ in range(args.warmup):
            ctx.render_direct(s, cam, W, H, LIGHTS, want_stats=False)
        device_sync(ctx)
        ctx.profile_begin()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.render_direct(s, cam, W, H, LIGHTS, want_stats=False)
        device_sync(ctx)
        wall = time.perf_counter() - t0
        prof = ctx.profile_end()
        st = ctx.render_direct(s, cam, W, H, LIGHTS, flags=rtcore.FLAG_COUNT)
        rays = prof["primary_rays"] + prof["shadow_rays"]
        per_launch_ms = prof["trace_ms"] / prof["launches"]
        per_rays = st["primary_rays"] + st["shadow_rays"]
        alg = RAY_BYTES * per_rays + BOX_BYTES * st["box_tests"] + TRI_BYTES * st["tri_tests"]
        rates[name] = {"mrays_s": round(rays / (prof["trace_ms"] * 1e-3) / 1e6, 2),
                       "trace_ms_per_frame": round(per_launch_ms, 4), "wall": wall,
                       "box_per_ray": round(st["box_tests"] / per_rays, 2),
                       "tri_per_ray": round(st["tri_tests"] / per_rays, 2), "tuning": list(tuned),
                       "achieved": alg / (per_launch_ms * 1e-3) / 1e9, "bytes_per_launch": alg}
    ctx.close()
    b = rates["bvh"]
    return result(METRIC, b["mrays_s"], "Mrays/s", args.steps, args.warmup, b.pop("wall"),
                  {"workload": f"{NTRI}-triangle soup (static), SAH BVH (1 triangle per leaf) against the "
                               f"LBVH (depth {DEPTH}), {W}x{H} primary + shadow rays, 1 point light",
                   "tris": NTRI, "width": W, "height": H},
                  bvh={"mrays_s": b["mrays_s"], "trace_ms_per_frame": b["trace_ms_per_frame"],
                       "box_per_ray": b["box_per_ray"], "tri_per_ray": b["tri_per_ray"],
                       "sah_cost": round(bvh_cost, 3), "build_ms_host_1thread": round(info["build_ms"], 2),
                       "max_depth": info["max_depth"], "forced_medians": info["forced_medians"],
                       "tuning": b["tuning"]},
                  lbvh={"mrays_s": rates["lbvh"]["mrays_s"], "trace_ms_per_frame": rates["lbvh"]["trace_ms_per_frame"],
                        "box_per_ray": rates["lbvh"]["box_per_ray"], "tri_per_ray": rates["lbvh"]["tri_per_ray"],
                        "sah_cost": round(lbvh_cost, 3), "build_ms_device": round(lb["total_ms"], 4),
                        "tuning": rates["lbvh"]["tuning"]},
                  roofline={"bound": "hbm", "achieved": round(b["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(b["achieved"] / HBM_PEAK_GBS, 4), "traffic": None,
                            "kernel": "k_render_direct", "bytes_per_launch": b["bytes_per_launch"]},
                  cpu_baseline=None)


# ------------------------------------------------------------------ batched
def gi_rmse_vs_fp64(ctx, rtcore, xyz, idx, depth, cam_fn, seed=7):
    """One scene at 48x48: 1024 device samples against 4096 independent FP64
    oracle samples per pixel; RMSE of the per-pixel mean radiance."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    w = h = 48
    s = ctx.add_scene(xyz, idx, depth)
    ctx.build(s)
    cam = cam_fn(w, h)
    acc = ctx.alloc(16 * w * h)
    ctx.zero(acc, 16 * w * h)
    ctx.render_jobs([ctx.job(s, cam, w, h, acc, rtcore.MODE_GI, 1024, 4, 0.7, 1.0, seed=seed)],
                    want_stats=False)
    g = ctx.read_accum(acc, w, h)[..., 0] / 1024
    ctx.free(acc)
    ctx.scene_destroy(s)
    ref = oracle.build_scene(xyz, idx, depth)
    r, _ = oracle.render_gi(ref["nodes"], ref["ids"], xyz, idx, cam.as_array(), w, h, 4096, 4, 0.7, 1.0,
                            seed, first_sample=1 << 20, nthreads=min(CPU_SHARE, len(os.sched_getaffinity(0))))
    return float(np.sqrt(np.mean((g - r / 4096) ** 2)))


def workload_batched(args, rtcore):
    n = 64
    rng = np.random.default_rng(44)
    sizes = np.exp(rng.uniform(math.log(1024), math.log(200000), n)).astype(np.int64)
    res = rng.choice([256, 512, 1024], n)
    scenes = [rtcore.gen_object_mix(int(sizes[k]), seed=1000 + k) for k in range(n)]
    depths = [depth_for(len(i)) for _, i in scenes]
    cams = [view_camera(rtcore, x, 0, int(r), int(r)) for (x, _), r in zip(scenes, res)]
    ctx = rtcore.Context(0)
    hs = ctx.batch_create([len(i) for _, i in scenes], [len(x) for x, _ in scenes], depths)
    for h_, (x, i) in zip(hs, scenes):
        ctx.scene_upload(h_, x, i)
    ps = [ctx.add_scene(x, i, d) for (x, i), d in zip(scenes, depths)]
    acc_b = [ctx.alloc(16 * int(r) * int(r)) for r in res]
    acc_p = [ctx.alloc(16 * int(r) * int(r)) for r in res]

    def jobs(handles, accs, frame):
        return [ctx.job(handles[k], cams[k], int(res[k]), int(res[k]), accs[k], rtcore.MODE_GI, 4, 4, 0.7, 1.0,
                        seed=7, frame=frame) for k in range(n)]

    ctx.build(hs[0])
    t0 = time.perf_counter()
    tuned = ctx.autotune(jobs(hs, acc_b, 0))
    tune_s = time.perf_counter() - t0

    def zero_all():
        for a, r in zip(acc_b + acc_p, list(res) + list(res)):
            ctx.zero(a, 16 * int(r) * int(r))

    def run_batched(frame0, steps):
        for f in range(frame0, frame0 + steps):
            ctx.build(hs[0])
            ctx.render_jobs(jobs(hs, acc_b, f), persistent=True, want_stats=False)

    def run_plain(frame0, steps):
        for f in range(frame0, frame0 + steps):
            js = jobs(ps, acc_p, f)
            for k in range(n):
                ctx.build(ps[k])
                ctx.render_jobs([js[k]], persistent=False, want_stats=False)

    def timed(fn):
        fn(0, args.warmup)
        ctx.sync()
        ctx.profile_begin()
        t0 = time.perf_counter()
        fn(args.warmup, args.steps)
        ctx.sync()
        wall = time.perf_counter() - t0
        return ctx.profile_end(), wall

    log(f"batched: autotuned {tuned} in {tune_s:.1f} s")
    zero_all()
    pb, wall_b = timed(run_batched)
    log(f"batched: {wall_b:.2f} s for {args.steps} steps")
    pp, wall_p = timed(run_plain)
    log(f"plain: {wall_p:.2f} s for {args.steps} steps")
    same = all(np.array_equal(ctx.read_accum(acc_b[k], int(res[k]), int(res[k])),
                              ctx.read_accum(acc_p[k], int(res[k]), int(res[k]))) for k in range(n))
    # the batched frame as a captured graph (rebuild + persistent launch),
    # replayed; frame indices advance per replay inside the graph
    g = ctx.graph_capture(hs[0], jobs(hs, acc_b, 0), rebuild=True)
    ctx.graph_launch(g)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.graph_launch(g)
    ctx.sync()
    wall_g = time.perf_counter() - t0
    ctx.graph_destroy(g)
    rays_b = pb["primary_rays"] + pb["bounce_rays"]
    rays_p = pp["primary_rays"] + pp["bounce_rays"]
    mr_b = rays_b / (pb["trace_ms"] * 1e-3) / 1e6
    mr_p = rays_p / (pp["trace_ms"] * 1e-3) / 1e6
    k_mid = int(np.argsort(sizes)[n // 2])
    rmse = None
    if not args.no_cpu:
        x, i = scenes[k_mid]
        rmse = gi_rmse_vs_fp64(ctx, rtcore, x, i, depths[k_mid],
                               lambda w, h: view_camera(rtcore, x, 0, w, h))
    ctx.close()
    return result(METRIC, mr_b, "Mrays/s", args.steps, args.warmup, wall_b,
                  {"workload": "64 object-mix scenes (1,024..200,000 tris, log-uniform), 256^2/512^2/1024^2, "
                               "GI 4 spp x 4 bounces, albedo 0.7; step = rebuild all + render all",
                   "scenes": n, "tris_total": int(sizes.sum()), "pixels_total": int((res * res).sum())},
                  plain={"mrays_s": round(mr_p, 2), "ms_per_step": round(wall_p * 1e3 / args.steps, 3),
                         "build_ms_per_step": round(pp["build_ms"]["total"] / args.steps, 4),
                         "trace_ms_per_step": round(pp["trace_ms"] / args.steps, 3),
                         "launches_per_step": pp["launches"] / args.steps},
                  batched={"mrays_s": round(mr_b, 2), "ms_per_step": round(wall_b * 1e3 / args.steps, 3),
                           "build_ms_per_step": round(pb["build_ms"]["total"] / args.steps, 4),
                           "trace_ms_per_step": round(pb["trace_ms"] / args.steps, 3),
                           "graph_ms_per_step": round(wall_g * 1e3 / args.steps, 3)},
                  speedup_trace=round(mr_b / mr_p, 3), speedup_step=round(wall_p / wall_b, 3),
                  identical_images=bool(same), tuning=list(tuned), autotune_s=round(tune_s, 2),
                  rays_per_step=rays_b / args.steps,
                  gi_rmse_vs_fp64={"value": rmse, "scene_tris": int(sizes[k_mid]), "res": "48x48",
                                   "gpu_spp": 1024, "fp64_spp": 4096}, cpu_baseline=None)


# ------------------------------------------------------------------ serve
# The server holds an arrived request up to this long, while nothing is in
# flight, so that later arrivals share its launch (rt_server_config.batch_window).
SERVE_WINDOW = 0.001


def workload_serve(args, rtcore):
    rng = np.random.default_rng(55)
    npool = 500
    sizes = np.exp(rng.uniform(math.log(1000), math.log(500000), npool)).astype(np.int64)
    duration, rate = args.duration, 200.0
    gaps = rng.exponential(1.0 / rate, int(rate * duration * 1.5))
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < duration]
    nreq = len(arrivals)
    scene_of = rng.integers(0, npool, nreq)
    big = rng.random(nreq) < 0.5
    gi = rng.random(nreq) < 0.5
    view_of = rng.integers(0, 4, nreq)
    pool = {}
    for k in np.unique(scene_of):
        pool[int(k)] = rtcore.gen_object_mix(int(sizes[k]), seed=5000 + int(k))
    log(f"serve: {nreq} requests over {len(pool)} scenes generated")

    def request(r):
        """(scene, view, w, h, mode, bounces, camera, lights) of request r."""
        k = int(scene_of[r])
        xyz = pool[k][0]
        w, h = (1920, 1080) if big[r] else (512, 512)
        cam = view_camera(rtcore, xyz, int(view_of[r]), w, h)
        if gi[r]:
            return k, int(view_of[r]), w, h, rtcore.MODE_GI, 2, cam, None
        light = tuple(xyz.max(0) + 0.5 * (xyz.max(0) - xyz.min(0)))
        return k, int(view_of[r]), w, h, rtcore.MODE_DIRECT, 0, cam, [light]

    def served(ctx, arr, window=SERVE_WINDOW):
        srv = rtcore.Server(ctx, 200 << 30, max_inflight=32, tiles_per_step=32768,
                            tile_order=rtcore.TILES_CENTRE, batch_window=window)
        for k, (x, i) in pool.items():
            srv.add_scene(k, x, i, depth_for(len(i)))
        for r, t in enumerate(arr):
            k, v, w, h, mode, nb, cam, lights = request(r)
            if mode == rtcore.MODE_GI:
                srv.submit(k, v, cam, w, h, mode, 1, nb, 0.7, 1.0, seed=3, arrival=float(t))
            else:
                srv.submit(k, v, cam, w, h, mode, 1, 0, 0.8, 0.0, 0.1, lights=lights, seed=3, arrival=float(t))
        t0 = time.perf_counter()
        res, t_end = srv.run(list(arr), progress=lambda t, st: log(
            f"serve: clock {t:.1f} s, {st['completed']} done, {st['pending']} pending"))
        wall = time.perf_counter() - t0
        st = srv.status()
        srv.close()
        ok = [r for r in res if r["status"] == rtcore.RT_OK]
        lat = np.array([r["completed"] - r["arrival"] for r in ok])
        return {"requests": len(arr), "completed": len(ok), "rejected": int(st["rejected"]),
                "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                "mrays_s_device": round(st["rays"] / st["device_s"] / 1e6, 2),
                "mrays_s_clock": round(st["rays"] / t_end / 1e6, 2),
                "device_busy_frac": round(st["device_s"] / t_end, 4), "clock_s": round(t_end, 3),
                "wall_s": round(wall, 2), "steps": int(st["steps"]), "builds": int(st["builds"]),
                "evictions": int(st["evictions"]), "rays": int(st["rays"]), "batch_window_ms": window * 1e3}

    def plain(ctx, arr):
        """The plain per-frame path: requests in arrival order, one at a time,
        each image rendered whole by its own launch (one block per tile),
        scenes built on first use and kept, one accumulator per stream."""
        scenes, accs = {}, {}
        t = dev = 0.0
        rays = 0
        lat = []
        for r, ta in enumerate(arr):
            k, v, w, h, mode, nb, cam, lights = request(r)
            t = max(t, float(ta))
            t0 = time.perf_counter()
            if k not in scenes:
                x, i = pool[k]
                scenes[k] = ctx.add_scene(x, i, depth_for(len(i)))
                ctx.build(scenes[k])
            key = (k, v, w, h, mode)
            if key not in accs:
                accs[key] = [ctx.alloc(16 * w * h), 0]
                ctx.zero(accs[key][0], 16 * w * h)
            acc = accs[key]
            if mode == rtcore.MODE_GI:
                job = ctx.job(scenes[k], cam, w, h, acc[0], mode, 1, nb, 0.7, 1.0, seed=3, frame=acc[1])
            else:
                job = ctx.job(scenes[k], cam, w, h, acc[0], mode, 1, 0, 0.8, 0.0, 0.1, lights=lights, seed=3,
                              frame=acc[1])
            st = ctx.render_jobs([job], persistent=False)
            acc[1] += 1
            t += time.perf_counter() - t0
            dev += st["trace_ms"] * 1e-3
            rays += st["primary_rays"] + st["shadow_rays"] + st["bounce_rays"]
            lat.append(t - float(ta))
        for a, _ in accs.values():
            ctx.free(a)
        for hnd in scenes.values():
            ctx.scene_destroy(hnd)
        lat = np.array(lat)
        return {"requests": len(arr), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                "mrays_s_device": round(rays / dev / 1e6, 2), "mrays_s_clock": round(rays / t / 1e6, 2),
                "device_busy_frac": round(dev / t, 4), "clock_s": round(t, 3), "rays": int(rays)}

    ctx = rtcore.Context(0)
    # warm-up: first launches of every kernel shape used below
    srv = rtcore.Server(ctx, 1 << 34, 4, 4096)
    k0 = int(scene_of[0])
    srv.add_scene(k0, *pool[k0], depth_for(len(pool[k0][1])))
    srv.submit(k0, 0, view_camera(rtcore, pool[k0][0], 0, 256, 256), 256, 256, arrival=0.0)
    srv.run([0.0])
    srv.close()
    s_poisson = served(ctx, arrivals)
    log(f"serve: continuous batching, Poisson arrivals: {s_poisson}")
    s_nowin = served(ctx, arrivals, window=0.0)
    log(f"serve: continuous batching without the batch window: {s_nowin}")
    nplain = min(nreq, int(rate * args.plain_seconds))
    p_poisson = plain(ctx, arrivals[:nplain])
    log(f"serve: plain path, Poisson arrivals: {p_poisson}")
    nburst = min(nreq, args.burst)
    s_burst = served(ctx, np.zeros(nburst))
    log(f"serve: continuous batching, burst: {s_burst}")
    p_burst = plain(ctx, np.zeros(nburst))
    log(f"serve: plain path, burst: {p_burst}")
    ctx.close()
    return result(METRIC, s_poisson["mrays_s_device"], "Mrays/s", nreq, 1, s_poisson["wall_s"],
                  {"workload": f"Poisson {rate:.0f} req/s for {duration:.0f} s, pool of {npool} object-mix "
                               "scenes (1k..500k tris, log-uniform), 512^2 or 1920x1080, direct (1 light) or "
                               "2-bounce GI at 1 spp, progressive accumulation per (scene, view, size, mode), "
                               "200 GB device budget",
                   "requests": nreq, "pool": npool},
                  p50_ms=s_poisson["p50_ms"], p99_ms=s_poisson["p99_ms"],
                  served=s_poisson, served_no_window=s_nowin, plain=p_poisson,
                  served_vs_plain_device=round(s_poisson["mrays_s_device"] / p_poisson["mrays_s_device"], 3),
                  burst={"requests": nburst, "served": s_burst, "plain": p_burst,
                         "speedup_clock": round(s_burst["mrays_s_clock"] / p_burst["mrays_s_clock"], 3),
                         "speedup_device": round(s_burst["mrays_s_device"] / p_burst["mrays_s_device"], 3)},
                  clock="virtual: each step (served) or request (plain) advances it by the measured wall "
                        "time of its library calls (admission, builds, synchronised launch); an idle "
                        "server jumps to the next arrival",
                  cpu_baseline=None)


def workload_all(args, rtcore):
    """The default line: the frame workload (value, roofline, CPU baseline),
    with the metric's other two parts beside it: the batched GI workload
    (64 scenes, batched against per-scene launches, images compared bit for
    bit, GI RMSE against the FP64 oracle) and the served stream (Poisson
    200/s for 60 s against one request per launch, p50/p99)."""
    import types

    out = workload_frame(args, rtcore)
    t0 = time.perf_counter()
    b = workload_batched(types.SimpleNamespace(steps=3, warmup=1, no_cpu=args.no_cpu), rtcore)
    out["batched"] = {"workload": b["config"]["workload"], "steps": 3, "mrays_s": b["value"],
                      "plain_mrays_s": b["plain"]["mrays_s"], "speedup_trace": b["speedup_trace"],
                      "speedup_step": b["speedup_step"], "ms_per_step": b["batched"]["ms_per_step"],
                      "plain_ms_per_step": b["plain"]["ms_per_step"],
                      "graph_ms_per_step": b["batched"]["graph_ms_per_step"],
                      "identical_images": b["identical_images"], "tuning": b["tuning"],
                      "gi_rmse_vs_fp64": b["gi_rmse_vs_fp64"], "wall_s": round(time.perf_counter() - t0, 1)}
    t0 = time.perf_counter()
    sv = workload_serve(types.SimpleNamespace(duration=60.0, plain_seconds=10.0, burst=500), rtcore)
    out["serve"] = {"workload": sv["config"]["workload"], "requests": sv["config"]["requests"],
                    "p50_ms": sv["p50_ms"], "p99_ms": sv["p99_ms"],
                    "served_mrays_s_device": sv["served"]["mrays_s_device"],
                    "plain_mrays_s_device": sv["plain"]["mrays_s_device"],
                    "served_vs_plain_device": sv["served_vs_plain_device"],
                    "batch_window_ms": SERVE_WINDOW * 1e3,
                    "no_window": {k: sv["served_no_window"][k] for k in ("p50_ms", "p99_ms", "mrays_s_device")},
                    "plain_p50_ms": sv["plain"]["p50_ms"], "plain_p99_ms": sv["plain"]["p99_ms"],
                    "device_busy_frac": sv["served"]["device_busy_frac"],
                    "burst_requests": sv["burst"]["requests"],
                    "burst_speedup_device": sv["burst"]["speedup_device"],
                    "burst_speedup_clock": sv["burst"]["speedup_clock"], "clock": sv["clock"],
                    "wall_s": round(time.perf_counter() - t0, 1)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", choices=["all", "frame", "sphere", "whitted", "bvh", "build", "batched", "serve"],
                    default="all")
    ap.add_argument("--autotune", action="store_true",
                    help="frame: pick the launch shape with rt_autotune_direct instead of FRAME_TUNING")
    ap.add_argument("--tris", type=int, default=0, help="build: one soup size instead of the list")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--plain-seconds", type=float, default=10.0,
                    help="serve: seconds of the same arrivals replayed one request per launch")
    ap.add_argument("--burst", type=int, default=2000,
                    help="serve: requests offered at once to compare saturated throughput")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the traversal shards by "
              f"image tile / scene across GPUs, which this build does not run", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py is single-process; WORLD_SIZE > 1 is not supported", file=sys.stderr)
        sys.exit(2)
    defaults = {"all": (100, 5), "frame": (100, 5), "sphere": (10, 1), "whitted": (20, 2), "bvh": (50, 3), "build": (20, 2),
                "batched": (3, 1), "serve": (0, 0)}
    ds, dw = defaults[args.workload]
    args.steps = ds if args.steps is None else args.steps
    args.warmup = dw if args.warmup is None else args.warmup

    # torch first: its HIP runtime is then the one librtcore.so binds to
    import torch

    import rtcore

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    rtcore.lib()
    fn = {"all": workload_all, "frame": workload_frame, "sphere": workload_sphere, "whitted": workload_whitted, "bvh": workload_bvh,
          "build": workload_build, "batched": workload_batched, "serve": workload_serve}[args.workload]
    print(json.dumps(fn(args, rtcore)), flush=True)


if __name__ == "__main__":
    main()
