"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the CPU baseline -- never by
the shipped library.  See lbvh_oracle.h for what each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

NODE_DTYPE = np.dtype([("mn", "<f4", 3), ("mx", "<f4", 3), ("prim_count", "<i4"), ("prim_pos", "<i4")])
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("refl", "<f4", 3), ("tmin", "<f4"),
                      ("tmax", "<f4"), ("tri", "<i4")])
HIT_DTYPE = np.dtype([("tri", "<i4"), ("u", "<f4"), ("v", "<f4"), ("t", "<f4"), ("visited", "<i4"),
                      ("root", "<i4")])
MATERIAL_DTYPE = np.dtype([("diffuse", "<f4", 4), ("ambient", "<f4", 4), ("specular", "<f4", 4),
                           ("shininess", "<f4"), ("alpha", "<f4"), ("id", "<u4"), ("pad", "<u4")])
INVALID = -2147483647


class WhittedParams(C.Structure):
    _fields_ = [("light", C.c_float * 3), ("bounces", C.c_uint32), ("switches", C.c_uint32),
                ("env_flag", C.c_int32), ("env", C.c_float * 3), ("frame", C.c_uint32),
                ("seed", C.c_uint64), ("ref_order", C.c_int32)]


def build_library(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        build_library()
        L = C.CDLL(LIB_PATH)
        P, U32, I32 = C.c_void_p, C.c_uint32, C.c_int
        L.orc_prim_bounds.argtypes = [P, P, U32, P, P, P, P]
        L.orc_morton_codes.argtypes = [P, U32, P, P]
        L.orc_morton_closed_form.argtypes = [U32, U32, U32]
        L.orc_morton_closed_form.restype = U32
        L.orc_lbvh_build.argtypes = [P, P, U32, U32, P, P, P]
        L.orc_lbvh_build.restype = I32
        L.orc_sort_pairs.argtypes = [P, P, U32]
        L.orc_trace.argtypes = [P, P, P, P, P, P, U32, I32, P, I32]
        L.orc_raygen_pinhole.argtypes = [P, U32, U32, P]
        L.orc_raygen_pinhole_jitter.argtypes = [P, U32, U32, P, P]
        L.orc_pcg_jitter.argtypes = [C.c_uint64, U32, U32, U32, U32, P]
        L.orc_pcg32_draws.argtypes = [C.c_uint64, C.c_uint64, U32, P]
        L.orc_raygen_world.argtypes = [P, U32, U32, P, P]
        L.orc_render_direct.argtypes = [P, P, P, P, P, P, U32, P, U32, P, P, P, I32]
        L.orc_render_gi.argtypes = [P, P, P, P, P, U32, U32, U32, U32, C.c_double, C.c_double,
                                    C.c_uint64, U32, U32, P, I32]
        L.orc_render_gi.restype = C.c_uint64
        L.orc_render_whitted.argtypes = [P, P, P, P, P, P, P, P, U32, U32, C.POINTER(WhittedParams), P, P,
                                         P, I32]
        L.orc_render_whitted.restype = C.c_uint64
        L.orc_trace_bvh.argtypes = [P, P, P, P, P, U32, I32, I32, P, I32]
        L.orc_trace_brute.argtypes = [P, P, U32, P, U32, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def prim_bounds(xyz: np.ndarray, idx: np.ndarray):
    xyz = np.ascontiguousarray(xyz, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    nt = idx.size // 3
    bmin = np.empty((nt, 3), np.float32)
    bmax = np.empty((nt, 3), np.float32)
    cent = np.empty((nt, 3), np.float32)
    scene = np.empty(6, np.float32)
    lib().orc_prim_bounds(_p(xyz), _p(idx), nt, _p(bmin), _p(bmax), _p(cent), _p(scene))
    return bmin, bmax, cent, scene


def morton_codes(cent: np.ndarray, scene: np.ndarray) -> np.ndarray:
    cent = np.ascontiguousarray(cent, np.float32)
    scene = np.ascontiguousarray(scene, np.float32)
    codes = np.empty(len(cent), np.uint32)
    lib().orc_morton_codes(_p(cent), len(cent), _p(scene), _p(codes))
    return codes


def morton_closed_form(qx: int, qy: int, qz: int) -> int:
    return int(lib().orc_morton_closed_form(qx, qy, qz))


def sort_pairs(codes: np.ndarray, ids: np.ndarray | None = None):
    codes = np.array(codes, np.uint32, copy=True)
    ids = np.arange(len(codes), dtype=np.int32) if ids is None else np.array(ids, np.int32, copy=True)
    lib().orc_sort_pairs(_p(codes), _p(ids), len(codes))
    return codes, ids


def lbvh_build(codes: np.ndarray, ids: np.ndarray, depth: int, bmin: np.ndarray, bmax: np.ndarray):
    """Runs build_recursive on copies of (codes, ids); returns
    (nodes, permuted codes, permuted ids, reachable node count)."""
    codes = np.array(codes, np.uint32, copy=True)
    ids = np.array(ids, np.int32, copy=True)
    nodes = np.empty(1 << depth, NODE_DTYPE)
    n = lib().orc_lbvh_build(_p(codes), _p(ids), len(codes), depth,
                             _p(np.ascontiguousarray(bmin, np.float32)),
                             _p(np.ascontiguousarray(bmax, np.float32)), _p(nodes))
    return nodes, codes, ids, n


def build_scene(xyz, idx, depth: int, presorted: bool = True):
    """Full CPU build of a scene: bounds, Morton, then build_recursive.
    presorted=True sorts (code, id) first (the GPU's id order); False runs the
    reference's in-place partition on the unsorted codes."""
    bmin, bmax, cent, scene = prim_bounds(xyz, idx)
    codes = morton_codes(cent, scene)
    ids = np.arange(len(codes), dtype=np.int32)
    if presorted:
        codes, ids = sort_pairs(codes, ids)
    nodes, pcodes, pids, reached = lbvh_build(codes, ids, depth, bmin, bmax)
    return dict(bmin=bmin, bmax=bmax, cent=cent, scene=scene, codes=pcodes, ids=pids,
                nodes=nodes, reached=reached, unsorted_codes=morton_codes(cent, scene))


def trace(nodes, ids, xyz, idx, rays, start=None, anyhit=False, nthreads=1):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.empty(len(rays), HIT_DTYPE)
    st = None if start is None else np.ascontiguousarray(start, np.int32)
    lib().orc_trace(_p(np.ascontiguousarray(nodes, NODE_DTYPE)), _p(np.ascontiguousarray(ids, np.int32)),
                    _p(np.ascontiguousarray(xyz, np.float32)), _p(np.ascontiguousarray(idx, np.uint32)),
                    _p(rays), _p(st), len(rays), 1 if anyhit else 0, _p(hits), nthreads)
    return hits


def trace_bvh(nodes, ordered, xyz, idx, rays, anyhit=False, ref_leaf_loop=True, nthreads=1):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.empty(len(rays), HIT_DTYPE)
    lib().orc_trace_bvh(_p(np.ascontiguousarray(nodes, NODE_DTYPE)), _p(np.ascontiguousarray(ordered, np.int32)),
                        _p(np.ascontiguousarray(xyz, np.float32)), _p(np.ascontiguousarray(idx, np.uint32)),
                        _p(rays), len(rays), 1 if anyhit else 0, 1 if ref_leaf_loop else 0, _p(hits), nthreads)
    return hits


def trace_brute(xyz, idx, rays):
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.empty(len(rays), HIT_DTYPE)
    idx = np.ascontiguousarray(idx, np.uint32)
    lib().orc_trace_brute(_p(np.ascontiguousarray(xyz, np.float32)), _p(idx), idx.size // 3, _p(rays), len(rays),
                          _p(hits))
    return hits


def raygen_pinhole(cam14: np.ndarray, w: int, h: int, jitter: np.ndarray | None = None) -> np.ndarray:
    """Pinhole primary rays; jitter: float32 [w*h, 2] sub-pixel offsets or None."""
    rays = np.empty(w * h, RAY_DTYPE)
    cam = _p(np.ascontiguousarray(cam14, np.float32))
    if jitter is None:
        lib().orc_raygen_pinhole(cam, w, h, _p(rays))
    else:
        lib().orc_raygen_pinhole_jitter(cam, w, h, _p(np.ascontiguousarray(jitter, np.float32)), _p(rays))
    return rays


def pcg32_draws(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint32)
    lib().orc_pcg32_draws(initstate, initseq, n, _p(out))
    return out


def pcg_jitter(seed: int, frame: int, sample: int, w: int, h: int) -> np.ndarray:
    """The render jobs' per-pixel jitter (jx, jy) of one sample: float32 [w*h, 2]."""
    out = np.empty((w * h, 2), np.float32)
    lib().orc_pcg_jitter(seed, frame, sample, w, h, _p(out))
    return out


def raygen_world(world16: np.ndarray, w: int, h: int, jitter: np.ndarray | None = None) -> np.ndarray:
    """Primary rays of the reference's camera matrix (g_mfWorld, row-major Matrix4)."""
    rays = np.empty(w * h, RAY_DTYPE)
    lib().orc_raygen_world(_p(np.ascontiguousarray(world16, np.float32).reshape(16)), w, h,
                           None if jitter is None else _p(np.ascontiguousarray(jitter, np.float32)),
                           _p(rays))
    return rays


def render_direct(nodes, ids, xyz, idx, rays, hits, lights, ambient=0.1, albedo=0.8, env=0.0,
                  nthreads=1):
    lights = np.ascontiguousarray(np.asarray(lights, np.float32).reshape(-1))
    nl = lights.size // 3
    n = len(rays)
    rgba = np.empty((n, 4), np.float32)
    occ = np.empty((n, max(nl, 1)), np.uint8)
    shade = np.array([ambient, albedo, env], np.float32)
    lib().orc_render_direct(_p(np.ascontiguousarray(nodes, NODE_DTYPE)),
                            _p(np.ascontiguousarray(ids, np.int32)),
                            _p(np.ascontiguousarray(xyz, np.float32)),
                            _p(np.ascontiguousarray(idx, np.uint32)),
                            _p(np.ascontiguousarray(rays, RAY_DTYPE)),
                            _p(np.ascontiguousarray(hits, HIT_DTYPE)), n, _p(lights), nl, _p(shade),
                            _p(rgba), _p(occ), nthreads)
    return rgba, occ


def render_direct_job(nodes, ids, xyz, idx, w, h, spp, lights, ambient=0.1, albedo=0.8, sky=1.0, seed=0,
                      frame=0, cam14=None, world16=None, nthreads=1):
    """What one direct-mode render job adds to a zeroed accumulator: per
    sample s, the jittered primary ray (pcg_jitter), the closest hit, the
    direct term with the reference's shadow rays (render_direct; a miss adds
    sky), summed in float32 in sample order.  Returns (accum [h, w, 4],
    shadow rays cast, primary hits)."""
    total = np.zeros(w * h, np.float32)
    shadow = 0
    nhit = 0
    for s in range(spp):
        jit = pcg_jitter(seed, frame, s, w, h)
        rays = raygen_world(world16, w, h, jit) if world16 is not None else raygen_pinhole(cam14, w, h, jit)
        hits = trace(nodes, ids, xyz, idx, rays, nthreads=nthreads)
        rgba, occ = render_direct(nodes, ids, xyz, idx, rays, hits, lights, ambient, albedo, sky,
                                  nthreads=nthreads)
        total = (total + rgba[:, 0]).astype(np.float32)
        shadow += int((occ[:, :max(len(np.asarray(lights).reshape(-1)) // 3, 0)] != 2).sum())
        nhit += int((hits["tri"] >= 0).sum())
    acc = np.empty((h, w, 4), np.float32)
    acc[..., 0] = acc[..., 1] = acc[..., 2] = total.reshape(h, w)
    acc[..., 3] = np.float32(spp)
    return acc, shadow, nhit


def render_gi(nodes, ids, xyz, idx, cam14, w, h, spp, max_bounces, albedo, sky, seed, frame=0,
              first_sample=0, nthreads=1):
    """FP64 path tracing; returns (per-pixel radiance sums [h, w], rays traced)."""
    out = np.empty(w * h, np.float64)
    rays = lib().orc_render_gi(_p(np.ascontiguousarray(nodes, NODE_DTYPE)),
                               _p(np.ascontiguousarray(ids, np.int32)),
                               _p(np.ascontiguousarray(xyz, np.float32)),
                               _p(np.ascontiguousarray(idx, np.uint32)),
                               _p(np.ascontiguousarray(cam14, np.float32)), w, h, spp, max_bounces,
                               albedo, sky, seed, frame, first_sample, _p(out), nthreads)
    return out.reshape(h, w), int(rays)


def render_whitted(nodes, ids, xyz, idx, cam14, w, h, light, bounces=2, switches=1, env_flag=-1,
                   env=(0.0, 0.0, 0.0), seed=1, frame=0, normals=None, mat_ids=None, materials=None,
                   ref_order=False, nthreads=1):
    """Whitted frame (whitted_oracle.c); returns (rgba [h*w, 4], first-pass hits, first-pass
    shadow results, rays traced)."""
    p = WhittedParams()
    p.light[:] = [float(x) for x in light]
    p.bounces, p.switches, p.env_flag = bounces, switches, env_flag
    p.env[:] = [float(x) for x in env]
    p.frame, p.seed, p.ref_order = frame, seed, 1 if ref_order else 0
    n = w * h
    rgba = np.empty((n, 4), np.float32)
    hits = np.empty(n, HIT_DTYPE)
    occ = np.empty(n, np.uint8)
    nrm = None if normals is None else np.ascontiguousarray(normals, np.float32)
    mid = None if mat_ids is None else np.ascontiguousarray(mat_ids, np.uint32)
    mats = None if materials is None else np.ascontiguousarray(materials, MATERIAL_DTYPE)
    rays = lib().orc_render_whitted(_p(np.ascontiguousarray(nodes, NODE_DTYPE)),
                                    _p(np.ascontiguousarray(ids, np.int32)),
                                    _p(np.ascontiguousarray(xyz, np.float32)),
                                    _p(np.ascontiguousarray(idx, np.uint32)), _p(nrm), _p(mid), _p(mats),
                                    _p(np.ascontiguousarray(cam14, np.float32)), w, h, C.byref(p),
                                    _p(rgba), _p(hits), _p(occ), nthreads)
    return rgba, hits, occ, int(rays)


def reachable(nodes: np.ndarray) -> np.ndarray:
    """Heap indices reachable from the root (node 1) through internal nodes."""
    out = []
    stack = [1]
    while stack:
        n = stack.pop()
        out.append(n)
        pc = int(nodes[n]["prim_count"])
        if not (pc & 4):
            stack.append(2 * n)
            stack.append(2 * n + 1)
    return np.array(sorted(out), np.int64)
