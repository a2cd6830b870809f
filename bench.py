"""Benchmark: per-frame LBVH rebuild + primary/shadow traversal on one GPU.

Workload (one step = one frame): 262,144 triangles (centroids U[0,1]^3, edges
N(0, 0.004^2)), rigidly rotated 1 degree per frame about the axis (1, 2, 3)
through (0.5, 0.5, 0.5); the LBVH is rebuilt every frame (bounds, Morton,
onesweep sort, emit, refit), then a 1920x1080 frame traces one primary ray per
pixel and one shadow ray per lit-side hit to a point light.

value = (primary + shadow rays) / summed device time of the trace kernel
(HIP events on the library's stream), the Mrays/s definition that excludes
the build; build_ms is the device time of one rebuild.  Scene data is resident
in device memory before the timed region; nothing is copied per frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "raytracer-gpupro4_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "Mrays/s + LBVH build ms @262k tris; batched vs plain; p50/p99 latency"

NTRI, SIGMA, SEED, DEPTH = 262144, 0.004, 2, 20
W, H = 1920, 1080
CAM_POS, CAM_TGT, FOV = (0.5, 0.5, -1.5), (0.5, 0.5, 0.5), 60.0
LIGHTS = [(0.5, 2.5, -1.0)]
AXIS, CENTER = (1.0, 2.0, 3.0), (0.5, 0.5, 0.5)
# algorithmic bytes per traced ray and per test (ray 48 B in, hit 24 B out,
# 32 B node box per box test, 4 B id + 12 B indices + 36 B vertices per
# triangle test)
RAY_BYTES, BOX_BYTES, TRI_BYTES = 48 + 24, 32, 52


def frame(ctx, rtcore, s, cam, f, flags=0, stats=False):
    ctx.scene_transform(s, rtcore.rotation_about(AXIS, float(f), CENTER))
    ctx.build(s)
    return ctx.render_direct(s, cam, W, H, LIGHTS, flags=flags, want_stats=stats)


def cpu_baseline(rtcore, xyz, idx, cam):
    """Oracle build (1 thread, as the reference builds on the CPU) and oracle
    traversal (all cores in the affinity mask) of a bounded sample of frame 0:
    every 8th pixel's primary ray plus its shadow ray (reference semantics)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle

    cores = len(os.sched_getaffinity(0))
    m = rtcore.rotation_about(AXIS, 0.0, CENTER)
    f = np.float32
    d = (xyz - m[9:12]).astype(f)
    R = m[:9].reshape(3, 3)
    rot = np.stack([((((R[r, 0] * d[:, 0]).astype(f) + (R[r, 1] * d[:, 1]).astype(f)).astype(f)
                      + (R[r, 2] * d[:, 2]).astype(f)).astype(f) + m[9 + r]).astype(f) for r in range(3)], 1)
    t0 = time.perf_counter()
    ref = oracle.build_scene(rot, idx, DEPTH, presorted=False)
    build_s = time.perf_counter() - t0
    rays = oracle.raygen_pinhole(cam.as_array(), W, H)[::8].copy()
    t0 = time.perf_counter()
    hits = oracle.trace(ref["nodes"], ref["ids"], rot, idx, rays, nthreads=cores)
    _, occ = oracle.render_direct(ref["nodes"], ref["ids"], rot, idx, rays, hits, LIGHTS, nthreads=cores)
    trace_s = time.perf_counter() - t0
    nrays = len(rays) + int((occ != 2).sum())
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": nrays / trace_s / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"frame 0, every 8th pixel of {W}x{H}: {len(rays)} primary + "
                      f"{nrays - len(rays)} shadow rays (closest hit from the hit leaf), "
                      f"oracle build of {NTRI} tris on 1 thread",
            "build_ms_1thread": build_s * 1e3, "trace_s": trace_s, "cpu_model": model}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the traversal shards by "
              f"image tile / scene across GPUs, which this build does not run", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py is single-process; WORLD_SIZE > 1 is not supported", file=sys.stderr)
        sys.exit(2)

    import torch

    import rtcore

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    rtcore.lib()
    xyz, idx = rtcore.gen_triangle_soup(NTRI, SIGMA, SEED)
    cam = rtcore.camera_look_at(CAM_POS, CAM_TGT, (0, 1, 0), FOV, W / H)
    ctx = rtcore.Context(0)
    s = ctx.add_scene(xyz, idx, DEPTH)
    for f in range(args.warmup):
        frame(ctx, rtcore, s, cam, f)
    ctx.sync()
    torch.cuda.synchronize()

    ctx.profile_begin()
    t0 = time.perf_counter()
    for f in range(args.steps):
        frame(ctx, rtcore, s, cam, f)
    ctx.sync()
    torch.cuda.synchronize()
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
    build = {k: v / prof["builds"] for k, v in prof["build_ms"].items()}

    out = {
        "metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{NTRI}-triangle soup rebuilt every frame (1 deg/frame rotation), "
                               f"{W}x{H} primary + shadow rays, 1 point light",
                   "tris": NTRI, "depth": DEPTH, "width": W, "height": H, "lights": len(LIGHTS)},
        "build_ms": round(build["total"], 4),
        "build_ms_stages": {k: round(v, 4) for k, v in build.items()},
        "trace_ms_per_frame": round(per_launch_ms, 4),
        "rays_per_frame": rays / args.steps,
        "shadow_rays_per_frame": prof["shadow_rays"] / args.steps,
        "frame_mrays_s_wall": round(rays / wall / 1e6, 2),
        "tests_per_ray": {"box": boxes / rays, "tri": tris / rays},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_render_direct", "bytes_per_launch": per_launch_bytes},
        "cpu_baseline": None,
    }
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(rtcore, xyz, idx, cam)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
