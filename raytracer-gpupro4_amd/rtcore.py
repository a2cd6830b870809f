"""ctypes binding of librtcore.so (the C ABI declared in include/rtcore.h).

This module is plumbing for tests, the bench and Python callers: it moves
numpy arrays across the C ABI and turns status codes into exceptions.  All
computation happens in the HIP kernels of librtcore.so; there is no Python or
CPU fallback, and importing fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librtcore.so")

RT_OK = 0
RT_ERR_ARG = -1
RT_ERR_HIP = -2
RT_ERR_NOMEM = -3
RT_ERR_STATE = -4
RT_ERR_CAPACITY = -5
RT_ERR_BUDGET = -6
RT_ERR_NODEV = -7

FLAG_ANYHIT = 1 << 0
FLAG_COUNT = 1 << 1
FLAG_KEEP_HITS = 1 << 2
FLAG_TIMING = 1 << 3
FLAG_PERSISTENT = 1 << 4
FLAG_START_AT_IROOT = 1 << 5
MODE_DIRECT = 0
MODE_GI = 1
NOALLOC_MAX_JOBS = 1024
TILES_ROWS = 0
TILES_CENTRE = 1
SERVER_TUNE_REQUESTS = 8
CAMERA_BASIS = 0
CAMERA_WORLD = 1
WHITTED_SHADOWS = 1 << 0
WHITTED_PHONG = 1 << 1
WHITTED_GLOSS = 1 << 2
WHITTED_ONE_PASS = 1 << 3
WHITTED_SPLIT = 1 << 4
WHITTED_QUADS = 1 << 5

# Byte layouts of RayTracerCS/Shaders/Structures.hlsl (Node 32 B, Ray 48 B,
# Intersection 24 B) as numpy dtypes.
NODE_DTYPE = np.dtype([("mn", "<f4", 3), ("mx", "<f4", 3), ("prim_count", "<i4"), ("prim_pos", "<i4")])
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("refl", "<f4", 3), ("tmin", "<f4"),
                      ("tmax", "<f4"), ("tri", "<i4")])
HIT_DTYPE = np.dtype([("tri", "<i4"), ("u", "<f4"), ("v", "<f4"), ("t", "<f4"), ("visited", "<i4"),
                      ("root", "<i4")])
# Material (Structures.hlsl:69-78, 64 B).
MATERIAL_DTYPE = np.dtype([("diffuse", "<f4", 4), ("ambient", "<f4", 4), ("specular", "<f4", 4),
                           ("shininess", "<f4"), ("alpha", "<f4"), ("id", "<u4"), ("pad", "<u4")])
assert NODE_DTYPE.itemsize == 32 and RAY_DTYPE.itemsize == 48 and HIT_DTYPE.itemsize == 24
assert MATERIAL_DTYPE.itemsize == 64


class BuildTimes(C.Structure):
    _fields_ = [("bounds_ms", C.c_float), ("morton_ms", C.c_float), ("sort_ms", C.c_float),
                ("emit_ms", C.c_float), ("refit_ms", C.c_float), ("total_ms", C.c_float)]

    def as_dict(self):
        return {k: float(getattr(self, k)) for k, _ in self._fields_}


class Profile(C.Structure):
    _fields_ = [("builds", C.c_uint32), ("build_ms", C.c_float * 6), ("launches", C.c_uint32),
                ("trace_ms", C.c_float), ("primary_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("bounce_rays", C.c_uint64), ("box_tests", C.c_uint64), ("tri_tests", C.c_uint64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["build_ms"] = dict(zip(("bounds", "morton", "sort", "emit", "refit", "total"),
                                 [float(x) for x in self.build_ms]))
        d["trace_ms"] = float(self.trace_ms)
        return d


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("fwd", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("tan_x", C.c_float), ("tan_y", C.c_float), ("kind", C.c_uint32),
                ("world", C.c_float * 16)]

    def as_array(self) -> np.ndarray:
        """The 14 floats {pos, fwd, right, up, tan_x, tan_y} of a basis camera."""
        return np.frombuffer(bytes(self), dtype=np.float32, count=14).copy()

    def world_matrix(self) -> np.ndarray:
        """The 16 floats of a CAMERA_WORLD camera's matrix (row-major Matrix4)."""
        return np.array(self.world[:], np.float32)


class RenderDesc(C.Structure):
    _fields_ = [("cam", Camera), ("width", C.c_uint32), ("height", C.c_uint32),
                ("lights", C.POINTER(C.c_float)), ("nlights", C.c_uint32), ("ambient", C.c_float),
                ("albedo", C.c_float), ("env", C.c_float)]


class WhittedDesc(C.Structure):
    _fields_ = [("cam", Camera), ("width", C.c_uint32), ("height", C.c_uint32), ("light", C.c_float * 3),
                ("bounces", C.c_uint32), ("switches", C.c_uint32), ("env_flag", C.c_int32),
                ("env", C.c_float * 3), ("frame", C.c_uint32), ("seed", C.c_uint64)]


class Model(C.Structure):
    _fields_ = [("xyz", C.c_void_p), ("normals", C.c_void_p), ("idx", C.c_void_p), ("mat_ids", C.c_void_p),
                ("materials", C.c_void_p), ("nv", C.c_uint32), ("nt", C.c_uint32), ("nmat", C.c_uint32)]


class BvhInfo(C.Structure):
    _fields_ = [("nodes", C.c_uint32), ("leaves", C.c_uint32), ("max_depth", C.c_uint32),
                ("forced_medians", C.c_uint32), ("sah_cost", C.c_float), ("build_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("block", C.c_int32), ("lds_stack", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("traversal", C.c_int32), ("quad_tiles", C.c_int32)]

    def as_tuple(self):
        return (self.block, self.lds_stack, self.blocks_per_cu, self.traversal, self.quad_tiles)


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_int), ("cam", Camera), ("width", C.c_uint32), ("height", C.c_uint32),
                ("mode", C.c_uint32), ("spp", C.c_uint32), ("max_bounces", C.c_uint32),
                ("albedo", C.c_float), ("sky", C.c_float), ("ambient", C.c_float),
                ("lights", C.POINTER(C.c_float)), ("nlights", C.c_uint32), ("seed", C.c_uint64),
                ("frame", C.c_uint32), ("accum", C.c_void_p), ("tile_begin", C.c_uint32),
                ("tile_count", C.c_uint32), ("tile_order", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_inflight", C.c_uint32), ("tiles_per_step", C.c_uint32),
                ("dry_run", C.c_int32), ("dry_rate", C.c_double), ("autotune", C.c_int32),
                ("tile_order", C.c_uint32), ("batch_window", C.c_double)]


class Request(C.Structure):
    _fields_ = [("scene_key", C.c_uint32), ("view", C.c_uint32), ("cam", Camera), ("width", C.c_uint32),
                ("height", C.c_uint32), ("mode", C.c_uint32), ("spp", C.c_uint32),
                ("max_bounces", C.c_uint32), ("albedo", C.c_float), ("sky", C.c_float),
                ("ambient", C.c_float), ("lights", C.POINTER(C.c_float)), ("nlights", C.c_uint32),
                ("seed", C.c_uint64), ("arrival", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("arrival", C.c_double), ("admitted", C.c_double),
                ("completed", C.c_double), ("rays", C.c_uint64), ("frame", C.c_uint32),
                ("samples", C.c_uint32), ("status", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("submitted", "completed", "rejected", "steps", "builds",
                                          "evictions", "rays", "bytes_used")] + \
               [(k, C.c_uint32) for k in ("pending", "inflight", "resident_scenes", "accumulators")] + \
               [("device_s", C.c_double), ("tune_s", C.c_double), ("tuning", Tuning), ("wake", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["tuning"] = self.tuning.as_tuple()
        return d


class FrameStats(C.Structure):
    _fields_ = [("trace_ms", C.c_float), ("primary_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("bounce_rays", C.c_uint64), ("box_tests", C.c_uint64), ("tri_tests", C.c_uint64)]

    def as_dict(self):
        return {k: (float(getattr(self, k)) if k == "trace_ms" else int(getattr(self, k)))
                for k, _ in self._fields_}


def build_library(force: bool = False) -> None:
    """Compiles librtcore.so in place with the package Makefile (hipcc, gfx950)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-j8"], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"librtcore.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P, U32, I32, U64 = C.c_void_p, C.c_uint32, C.c_int, C.c_uint64
    sig = {
        "rt_create": (I32, [C.POINTER(P), I32]),
        "rt_destroy": (None, [P]),
        "rt_last_error": (C.c_char_p, [P]),
        "rt_sync": (I32, [P]),
        "rt_get_stream": (I32, [P, C.POINTER(P)]),
        "rt_memory_in_use": (I32, [P, C.POINTER(U64)]),
        "rt_scene_create": (I32, [P, U32, U32, U32, C.POINTER(I32)]),
        "rt_scene_destroy": (I32, [P, I32]),
        "rt_batch_create": (I32, [P, U32, P, P, P, C.POINTER(I32)]),
        "rt_scene_upload": (I32, [P, I32, P, U32, P, U32]),
        "rt_scene_upload_device": (I32, [P, I32, P, U32, P, U32]),
        "rt_scene_transform": (I32, [P, I32, P]),
        "rt_scene_info": (I32, [P, I32, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32)]),
        "rt_lbvh_build": (I32, [P, I32, C.POINTER(BuildTimes)]),
        "rt_lbvh_read": (I32, [P, I32, P, P, P, P]),
        "rt_trace": (I32, [P, I32, P, U32, P, U32, C.POINTER(FrameStats)]),
        "rt_trace_device": (I32, [P, I32, P, U32, P, U32, C.POINTER(FrameStats)]),
        "rt_render_direct": (I32, [P, I32, C.POINTER(RenderDesc), U32, C.POINTER(FrameStats)]),
        "rt_frame_read": (I32, [P, P, P, P]),
        "rt_scene_shading": (I32, [P, I32, P, U32, P, U32, P, U32]),
        "rt_render_whitted": (I32, [P, I32, C.POINTER(WhittedDesc), U32, C.POINTER(FrameStats)]),
        "rt_autotune_whitted": (I32, [P, I32, C.POINTER(WhittedDesc), C.c_char_p, C.POINTER(Tuning)]),
        "rt_load_model_ex": (I32, [C.c_char_p, U32, C.POINTER(Model)]),
        "rt_present": (I32, [P, P, U32, U32, U32, P]),
        "rt_bvh_build": (I32, [P, I32, U32, U32, C.POINTER(BvhInfo)]),
        "rt_scene_select": (I32, [P, I32, I32]),
        "rt_bvh_read": (I32, [P, I32, P, P]),
        "rt_bvh_build_host": (I32, [P, U32, P, U32, U32, U32, P, P, C.POINTER(BvhInfo)]),
        "rt_image_write": (I32, [C.c_char_p, P, U32, U32]),
        "rt_image_last_error": (C.c_char_p, []),
        "rt_profile_begin": (I32, [P]),
        "rt_profile_end": (I32, [P, C.POINTER(Profile)]),
        "rt_tuning_get": (I32, [P, C.POINTER(Tuning)]),
        "rt_tuning_set": (I32, [P, C.POINTER(Tuning)]),
        "rt_render_jobs": (I32, [P, P, U32, U32, C.POINTER(FrameStats)]),
        "rt_render_jobs_noalloc": (I32, [P, P, U32, U32, C.POINTER(FrameStats)]),
        "rt_autotune_jobs": (I32, [P, P, U32, C.c_char_p, C.POINTER(Tuning)]),
        "rt_autotune_direct": (I32, [P, I32, C.POINTER(RenderDesc), C.c_char_p, C.POINTER(Tuning)]),
        "rt_graph_capture_frame": (I32, [P, I32, I32, P, U32, C.POINTER(I32)]),
        "rt_graph_launch": (I32, [P, I32]),
        "rt_graph_destroy": (I32, [P, I32]),
        "rt_device_alloc": (I32, [P, U64, C.POINTER(P)]),
        "rt_device_free": (I32, [P, P]),
        "rt_device_zero": (I32, [P, P, U64]),
        "rt_device_read": (I32, [P, P, P, U64]),
        "rt_server_create": (I32, [P, C.POINTER(ServerConfig), C.POINTER(P)]),
        "rt_server_destroy": (None, [P]),
        "rt_server_last_error": (C.c_char_p, [P]),
        "rt_server_add_scene": (I32, [P, U32, P, U32, P, U32, U32]),
        "rt_server_submit": (I32, [P, C.POINTER(Request), C.POINTER(U64)]),
        "rt_server_step": (I32, [P, C.c_double, C.POINTER(C.c_double)]),
        "rt_server_poll": (I32, [P, P, U32, C.POINTER(U32)]),
        "rt_server_status": (I32, [P, C.POINTER(ServerStats)]),
        "rt_server_read_stream": (I32, [P, U32, U32, U32, U32, U32, P, C.POINTER(U32)]),
        "rt_load_model": (I32, [C.c_char_p, U32, P, C.POINTER(U32), P, C.POINTER(U32)]),
        "rt_load_last_error": (C.c_char_p, []),
        "rt_camera_look_at": (None, [C.POINTER(Camera), P, P, P, C.c_float, C.c_float]),
        "rt_camera_from_world": (None, [C.POINTER(Camera), P]),
        "rt_gen_torus": (I32, [C.c_float, C.c_float, U32, U32, C.c_float, U64, P, C.POINTER(U32), P,
                               C.POINTER(U32)]),
        "rt_gen_triangle_soup": (I32, [U32, C.c_float, U64, P, C.POINTER(U32), P, C.POINTER(U32)]),
        "rt_gen_noisy_sphere": (I32, [U32, U32, C.c_float, U64, P, C.POINTER(U32), P, C.POINTER(U32)]),
        "rt_gen_object_mix": (I32, [U32, U64, P, C.POINTER(U32), P, C.POINTER(U32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class RtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (status {code})")
        self.code = code


# ---------------------------------------------------------------- generators
def _gen(fn, *args):
    L = lib()
    nv, nt = C.c_uint32(0), C.c_uint32(0)
    rc = getattr(L, fn)(*args, None, C.byref(nv), None, C.byref(nt))
    if rc != RT_OK:
        raise RtError(rc, fn)
    xyz = np.empty((nv.value, 3), np.float32)
    idx = np.empty((nt.value, 3), np.uint32)
    rc = getattr(L, fn)(*args, _ptr(xyz), C.byref(nv), _ptr(idx), C.byref(nt))
    if rc != RT_OK:
        raise RtError(rc, fn)
    return xyz, idx


LOAD_RAW = 1 << 0


def load_model(path, raw=False):
    """Vertices and triangles of a .obj or the reference's .txt model file,
    normalised as the reference normalises its models unless raw."""
    L = lib()
    nv, nt = C.c_uint32(0), C.c_uint32(0)
    flags = LOAD_RAW if raw else 0
    rc = L.rt_load_model(os.fsencode(path), flags, None, C.byref(nv), None, C.byref(nt))
    if rc != RT_OK:
        raise RtError(rc, L.rt_load_last_error().decode())
    xyz = np.empty((nv.value, 3), np.float32)
    idx = np.empty((nt.value, 3), np.uint32)
    rc = L.rt_load_model(os.fsencode(path), flags, _ptr(xyz), C.byref(nv), _ptr(idx), C.byref(nt))
    if rc != RT_OK:
        raise RtError(rc, L.rt_load_last_error().decode())
    return xyz, idx


def load_model_shading(path, raw=False) -> dict:
    """load_model plus shading attributes: dict with xyz, idx, normals (per
    vertex), mat_ids (per triangle) and materials (MATERIAL_DTYPE)."""
    L = lib()
    flags = LOAD_RAW if raw else 0
    m = Model()
    rc = L.rt_load_model_ex(os.fsencode(path), flags, C.byref(m))
    if rc != RT_OK:
        raise RtError(rc, L.rt_load_last_error().decode())
    out = dict(xyz=np.empty((m.nv, 3), np.float32), normals=np.empty((m.nv, 3), np.float32),
               idx=np.empty((m.nt, 3), np.uint32), mat_ids=np.empty(m.nt, np.uint32),
               materials=np.empty(m.nmat, MATERIAL_DTYPE))
    m.xyz, m.normals, m.idx = out["xyz"].ctypes.data, out["normals"].ctypes.data, out["idx"].ctypes.data
    m.mat_ids, m.materials = out["mat_ids"].ctypes.data, out["materials"].ctypes.data
    rc = L.rt_load_model_ex(os.fsencode(path), flags, C.byref(m))
    if rc != RT_OK:
        raise RtError(rc, L.rt_load_last_error().decode())
    return out


PRESENT_AVERAGE = 1 << 0
BVH_SAH, BVH_MIDDLE, BVH_EQUAL = 0, 1, 2
STRUCTURE_BVH, STRUCTURE_LBVH = 0, 9


def bvh_build_host(xyz, idx, max_prims=1, split=BVH_SAH):
    """The SAH BVH of the arrays on the host: (nodes NODE_DTYPE, ordered ids, info)."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    nt = idx.size // 3
    nodes = np.zeros(max(2 * nt - 1, 1), NODE_DTYPE)
    ids = np.empty(nt, np.int32)
    info = BvhInfo()
    rc = lib().rt_bvh_build_host(_ptr(xyz), xyz.size // 3, _ptr(idx), nt, max_prims, split, _ptr(nodes),
                                 _ptr(ids), C.byref(info))
    if rc != RT_OK:
        raise RtError(rc, "rt_bvh_build_host")
    return nodes[:info.nodes].copy(), ids, info.as_dict()


def image_write(path, rgba8: np.ndarray):
    """Writes an (h, w, 4) uint8 image as .png or .ppm (host only)."""
    img = np.ascontiguousarray(rgba8, np.uint8)
    h, w = img.shape[:2]
    L = lib()
    rc = L.rt_image_write(os.fsencode(path), _ptr(img), w, h)
    if rc != RT_OK:
        raise RtError(rc, L.rt_image_last_error().decode())


def gen_torus(R=1.0, r=0.3, nu=128, nv=64, jitter=1e-3, seed=1):
    return _gen("rt_gen_torus", R, r, nu, nv, jitter, seed)


def gen_triangle_soup(ntri=262144, sigma=0.004, seed=2):
    return _gen("rt_gen_triangle_soup", ntri, sigma, seed)


def gen_noisy_sphere(nlat=1025, nlon=1024, amplitude=0.05, seed=3):
    return _gen("rt_gen_noisy_sphere", nlat, nlon, amplitude, seed)


def gen_object_mix(target_tris, seed=4):
    return _gen("rt_gen_object_mix", target_tris, seed)


def camera_look_at(pos, target, up=(0.0, 1.0, 0.0), fov_y_deg=60.0, aspect=1.0) -> Camera:
    cam = Camera()
    p = np.asarray(pos, np.float32)
    t = np.asarray(target, np.float32)
    u = np.asarray(up, np.float32)
    lib().rt_camera_look_at(C.byref(cam), _ptr(p), _ptr(t), _ptr(u), fov_y_deg, aspect)
    return cam


def camera_from_world(world) -> Camera:
    """The reference's camera: its Matrix4 (cbCamera g_mfWorld), 16 floats row-major."""
    cam = Camera()
    m = np.ascontiguousarray(np.asarray(world, np.float32).reshape(16))
    lib().rt_camera_from_world(C.byref(cam), _ptr(m))
    return cam


def rotation_about(axis, degrees, center) -> np.ndarray:
    """12 floats {R row-major, c} for rt_scene_transform (rotation about an
    axis through `center`), computed in float64 and rounded once."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    th = np.deg2rad(degrees)
    x, y, z = a
    c, s = np.cos(th), np.sin(th)
    R = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                  [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                  [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])
    return np.concatenate([R.reshape(-1), np.asarray(center, np.float64)]).astype(np.float32)


# ------------------------------------------------------------------- context
class Context:
    """One HIP device + stream (rt_create)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        rc = L.rt_create(C.byref(h), device)
        if rc != RT_OK:
            raise RtError(rc, "rt_create")
        self.h = h
        self.L = L

    def close(self):
        if self.h:
            self.L.rt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int, what: str):
        if rc != RT_OK:
            raise RtError(rc, f"{what}: {self.L.rt_last_error(self.h).decode()}")

    def sync(self):
        self.check(self.L.rt_sync(self.h), "rt_sync")

    def stream(self) -> int:
        s = C.c_void_p()
        self.check(self.L.rt_get_stream(self.h, C.byref(s)), "rt_get_stream")
        return s.value or 0

    def memory_in_use(self) -> int:
        b = C.c_uint64()
        self.check(self.L.rt_memory_in_use(self.h, C.byref(b)), "rt_memory_in_use")
        return b.value

    def profile_begin(self):
        self.check(self.L.rt_profile_begin(self.h), "rt_profile_begin")

    def profile_end(self) -> dict:
        p = Profile()
        self.check(self.L.rt_profile_end(self.h, C.byref(p)), "rt_profile_end")
        return p.as_dict()

    # tuning, jobs, graphs, device buffers
    def tuning(self):
        t = Tuning()
        self.check(self.L.rt_tuning_get(self.h, C.byref(t)), "rt_tuning_get")
        return t.as_tuple()

    def set_tuning(self, block, lds_stack, blocks_per_cu=4, traversal=0, quad_tiles=0):
        t = Tuning(block, lds_stack, blocks_per_cu, traversal, quad_tiles)
        self.check(self.L.rt_tuning_set(self.h, C.byref(t)), "rt_tuning_set")

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self.check(self.L.rt_device_alloc(self.h, nbytes, C.byref(p)), "rt_device_alloc")
        return p.value

    def free(self, ptr: int):
        self.check(self.L.rt_device_free(self.h, C.c_void_p(ptr)), "rt_device_free")

    def zero(self, ptr: int, nbytes: int):
        self.check(self.L.rt_device_zero(self.h, C.c_void_p(ptr), nbytes), "rt_device_zero")

    def read_accum(self, ptr: int, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), np.float32)
        self.check(self.L.rt_device_read(self.h, _ptr(out), C.c_void_p(ptr), out.nbytes), "rt_device_read")
        return out

    @staticmethod
    def job(scene, cam, width, height, accum, mode=MODE_GI, spp=1, max_bounces=4, albedo=0.7,
            sky=1.0, ambient=0.1, lights=None, seed=1, frame=0, tile_begin=0, tile_count=0, tile_order=0):
        """Builds a JobDesc; keeps the light array alive on the object."""
        j = JobDesc()
        j.scene, j.cam, j.width, j.height = scene, cam, width, height
        j.mode, j.spp, j.max_bounces = mode, spp, max_bounces
        j.albedo, j.sky, j.ambient = albedo, sky, ambient
        arr = np.ascontiguousarray(np.asarray(lights if lights is not None else [], np.float32).reshape(-1))
        j._lights = arr
        j.lights = arr.ctypes.data_as(C.POINTER(C.c_float)) if arr.size else None
        j.nlights = arr.size // 3
        j.seed, j.frame, j.accum = seed, frame, accum
        j.tile_begin, j.tile_count, j.tile_order = tile_begin, tile_count, tile_order
        return j

    def render_jobs(self, jobs, persistent=True, want_stats=True, flags=0):
        """flags: FLAG_COUNT keeps one lane per ray and fills box_tests / tri_tests."""
        arr = (JobDesc * len(jobs))(*jobs)
        st = FrameStats()
        flags = flags | (FLAG_PERSISTENT if persistent else 0)
        self.check(self.L.rt_render_jobs(self.h, arr, len(jobs), flags,
                                         C.byref(st) if want_stats else None), "rt_render_jobs")
        return st.as_dict() if want_stats else None

    def render_jobs_noalloc(self, jobs, persistent=True, want_stats=True, flags=0):
        """rt_render_jobs_noalloc: the same render with no device allocation."""
        arr = (JobDesc * len(jobs))(*jobs)
        st = FrameStats()
        flags = flags | (FLAG_PERSISTENT if persistent else 0)
        self.check(self.L.rt_render_jobs_noalloc(self.h, arr, len(jobs), flags,
                                                 C.byref(st) if want_stats else None), "rt_render_jobs_noalloc")
        return st.as_dict() if want_stats else None

    def autotune(self, jobs, cache_path=None):
        arr = (JobDesc * len(jobs))(*jobs)
        t = Tuning()
        cp = cache_path.encode() if cache_path else None
        self.check(self.L.rt_autotune_jobs(self.h, arr, len(jobs), cp, C.byref(t)), "rt_autotune_jobs")
        return t.as_tuple()

    def autotune_direct(self, scene, cam, width, height, lights, ambient=0.1, albedo=0.8, env=0.0,
                        cache_path=None):
        """Picks the fastest launch shape for rt_render_direct of this frame."""
        lights = np.ascontiguousarray(np.asarray(lights, np.float32).reshape(-1))
        d = RenderDesc()
        d.cam = cam
        d.width, d.height = width, height
        d.lights = lights.ctypes.data_as(C.POINTER(C.c_float))
        d.nlights = lights.size // 3
        d.ambient, d.albedo, d.env = ambient, albedo, env
        t = Tuning()
        cp = cache_path.encode() if cache_path else None
        self.check(self.L.rt_autotune_direct(self.h, scene, C.byref(d), cp, C.byref(t)), "rt_autotune_direct")
        return t.as_tuple()

    def graph_capture(self, scene, jobs, rebuild=True) -> int:
        arr = (JobDesc * len(jobs))(*jobs)
        g = C.c_int()
        self.check(self.L.rt_graph_capture_frame(self.h, scene, 1 if rebuild else 0, arr, len(jobs),
                                                 C.byref(g)), "rt_graph_capture_frame")
        return g.value

    def graph_launch(self, g: int):
        self.check(self.L.rt_graph_launch(self.h, g), "rt_graph_launch")

    def graph_destroy(self, g: int):
        self.check(self.L.rt_graph_destroy(self.h, g), "rt_graph_destroy")

    # scenes
    def scene_create(self, max_tris: int, max_verts: int, depth: int) -> int:
        s = C.c_int()
        self.check(self.L.rt_scene_create(self.h, max_tris, max_verts, depth, C.byref(s)),
                   "rt_scene_create")
        return s.value

    def batch_create(self, max_tris, max_verts, depths) -> list:
        """Creates a batch of scenes built together; returns their handles."""
        mt = np.ascontiguousarray(max_tris, np.uint32)
        mv = np.ascontiguousarray(max_verts, np.uint32)
        dp = np.ascontiguousarray(depths, np.uint32)
        first = C.c_int()
        self.check(self.L.rt_batch_create(self.h, len(mt), _ptr(mt), _ptr(mv), _ptr(dp),
                                          C.byref(first)), "rt_batch_create")
        return list(range(first.value, first.value + len(mt)))

    def scene_destroy(self, scene: int):
        self.check(self.L.rt_scene_destroy(self.h, scene), "rt_scene_destroy")

    def scene_upload(self, scene: int, xyz: np.ndarray, idx: np.ndarray):
        xyz = np.ascontiguousarray(xyz, np.float32)
        idx = np.ascontiguousarray(idx, np.uint32)
        self.check(self.L.rt_scene_upload(self.h, scene, _ptr(xyz), xyz.size // 3, _ptr(idx),
                                          idx.size // 3), "rt_scene_upload")

    def scene_upload_device(self, scene: int, d_xyz: int, nv: int, d_idx: int, nt: int):
        self.check(self.L.rt_scene_upload_device(self.h, scene, C.c_void_p(d_xyz), nv,
                                                 C.c_void_p(d_idx), nt), "rt_scene_upload_device")

    def scene_transform(self, scene: int, m12: np.ndarray):
        m = np.ascontiguousarray(m12, np.float32)
        assert m.size == 12
        self.check(self.L.rt_scene_transform(self.h, scene, _ptr(m)), "rt_scene_transform")

    def scene_info(self, scene: int):
        nt, nv, d = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self.check(self.L.rt_scene_info(self.h, scene, C.byref(nt), C.byref(nv), C.byref(d)),
                   "rt_scene_info")
        return nt.value, nv.value, d.value

    def add_scene(self, xyz: np.ndarray, idx: np.ndarray, depth: int) -> int:
        s = self.scene_create(len(idx), len(xyz), depth)
        self.scene_upload(s, xyz, idx)
        return s

    # build
    def build(self, scene: int, timed: bool = False):
        t = BuildTimes()
        self.check(self.L.rt_lbvh_build(self.h, scene, C.byref(t) if timed else None), "rt_lbvh_build")
        return t.as_dict() if timed else None

    def read(self, scene: int):
        nt, _, depth = self.scene_info(scene)
        nodes = np.empty(1 << depth, NODE_DTYPE)
        codes = np.empty(nt, np.uint32)
        ids = np.empty(nt, np.int32)
        box = np.empty(6, np.float32)
        self.check(self.L.rt_lbvh_read(self.h, scene, _ptr(nodes), _ptr(codes), _ptr(ids), _ptr(box)),
                   "rt_lbvh_read")
        return nodes, codes, ids, box

    # trace
    def trace(self, scene: int, rays: np.ndarray, flags: int = 0, want_stats: bool = False, start=None):
        """Closest (or any, FLAG_ANYHIT) hit per ray.  start: per-ray heap node
        to begin at (the previous hits' "root", LBVH_IntersectP's iRoot); sets
        FLAG_START_AT_IROOT."""
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.empty(len(rays), HIT_DTYPE)
        if start is not None:
            hits["root"] = np.broadcast_to(np.asarray(start, np.int32), len(rays))
            flags |= FLAG_START_AT_IROOT
        st = FrameStats()
        self.check(self.L.rt_trace(self.h, scene, _ptr(rays), len(rays), _ptr(hits), flags,
                                   C.byref(st) if want_stats else None), "rt_trace")
        return (hits, st.as_dict()) if want_stats else hits

    def trace_device(self, scene: int, d_rays: int, n: int, d_hits: int, flags: int = 0,
                     want_stats: bool = False):
        st = FrameStats()
        self.check(self.L.rt_trace_device(self.h, scene, C.c_void_p(d_rays), n, C.c_void_p(d_hits),
                                          flags, C.byref(st) if want_stats else None),
                   "rt_trace_device")
        return st.as_dict() if want_stats else None

    def render_direct(self, scene: int, cam: Camera, width: int, height: int, lights,
                      ambient=0.1, albedo=0.8, env=0.0, flags: int = 0, want_stats: bool = True):
        lights = np.ascontiguousarray(np.asarray(lights, np.float32).reshape(-1))
        d = RenderDesc()
        d.cam = cam
        d.width, d.height = width, height
        d.lights = lights.ctypes.data_as(C.POINTER(C.c_float))
        d.nlights = lights.size // 3
        d.ambient, d.albedo, d.env = ambient, albedo, env
        st = FrameStats()
        self.check(self.L.rt_render_direct(self.h, scene, C.byref(d), flags,
                                           C.byref(st) if want_stats else None), "rt_render_direct")
        return st.as_dict() if want_stats else None

    def scene_shading(self, scene: int, normals=None, mat_ids=None, materials=None):
        nt, nv, _ = self.scene_info(scene)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32)
        mid = None if mat_ids is None else np.ascontiguousarray(mat_ids, np.uint32)
        mats = None if materials is None else np.ascontiguousarray(materials, MATERIAL_DTYPE)
        self.check(self.L.rt_scene_shading(self.h, scene, _ptr(nrm), nv, _ptr(mid), nt, _ptr(mats),
                                           0 if mats is None else len(mats)), "rt_scene_shading")

    @staticmethod
    def _whitted_desc(cam, width, height, light, bounces, switches, env_flag, env, seed, frame):
        d = WhittedDesc()
        d.cam = cam
        d.width, d.height = width, height
        d.light[:] = [float(x) for x in light]
        d.bounces, d.switches, d.env_flag = bounces, switches, env_flag
        d.env[:] = [float(x) for x in env]
        d.frame, d.seed = frame, seed
        return d

    def autotune_whitted(self, scene: int, cam: Camera, width: int, height: int, light, bounces=2,
                         switches=WHITTED_SHADOWS, env_flag=-1, env=(0.0, 0.0, 0.0), seed=1, frame=0,
                         cache_path=None):
        """Picks the launch shape and Whitted form; returns (tuning tuple, switches to use)."""
        d = self._whitted_desc(cam, width, height, light, bounces, switches, env_flag, env, seed, frame)
        t = Tuning()
        self.check(self.L.rt_autotune_whitted(self.h, scene, C.byref(d),
                                              cache_path.encode() if cache_path else None, C.byref(t)),
                   "rt_autotune_whitted")
        return t.as_tuple(), int(d.switches)

    def render_whitted(self, scene: int, cam: Camera, width: int, height: int, light, bounces=2,
                       switches=WHITTED_SHADOWS, env_flag=-1, env=(0.0, 0.0, 0.0), seed=1, frame=0,
                       flags: int = 0, want_stats: bool = True):
        d = self._whitted_desc(cam, width, height, light, bounces, switches, env_flag, env, seed, frame)
        st = FrameStats()
        self.check(self.L.rt_render_whitted(self.h, scene, C.byref(d), flags,
                                            C.byref(st) if want_stats else None), "rt_render_whitted")
        return st.as_dict() if want_stats else None

    def bvh_build(self, scene: int, max_prims: int = 1, split: int = BVH_SAH) -> dict:
        info = BvhInfo()
        self.check(self.L.rt_bvh_build(self.h, scene, max_prims, split, C.byref(info)), "rt_bvh_build")
        return info.as_dict()

    def scene_select(self, scene: int, structure: int):
        self.check(self.L.rt_scene_select(self.h, scene, structure), "rt_scene_select")

    def bvh_read(self, scene: int, nodes: int):
        nt = self.scene_info(scene)[0]
        out = np.empty(nodes, NODE_DTYPE)
        ids = np.empty(nt, np.int32)
        self.check(self.L.rt_bvh_read(self.h, scene, _ptr(out), _ptr(ids)), "rt_bvh_read")
        return out, ids

    def present(self, width: int, height: int, d_rgba: int | None = None, average: bool = False) -> np.ndarray:
        """(h, w, 4) uint8 image of the context's frame buffer, or of a device
        float RGBA image (an accumulation buffer with average=True)."""
        out = np.empty((height, width, 4), np.uint8)
        self.check(self.L.rt_present(self.h, None if d_rgba is None else C.c_void_p(d_rgba), width, height,
                                     PRESENT_AVERAGE if average else 0, _ptr(out)), "rt_present")
        return out

    def frame_read(self, width: int, height: int, nlights: int, hits: bool = False):
        rgba = np.empty((height * width, 4), np.float32)
        h = np.empty(height * width, HIT_DTYPE) if hits else None
        occ = np.empty((height * width, max(nlights, 1)), np.uint8) if hits else None
        self.check(self.L.rt_frame_read(self.h, _ptr(rgba), _ptr(h), _ptr(occ)), "rt_frame_read")
        return rgba, h, occ


class Server:
    """Render server (rt_server_*): resident scenes and progressive streams
    under a device-memory budget, continuous batching of requests.
    ctx=None requires dry_run (scheduling only, no device work)."""

    def __init__(self, ctx: Context | None, hbm_budget: int, max_inflight: int = 16,
                 tiles_per_step: int = 4096, dry_run: bool = False, dry_rate: float = 1e9,
                 autotune: bool = False, tile_order: int = TILES_ROWS, batch_window: float = 0.0):
        self.L = lib()
        self.ctx = ctx
        cfg = ServerConfig(hbm_budget, max_inflight, tiles_per_step, 1 if dry_run else 0, dry_rate,
                           1 if autotune else 0, tile_order, batch_window)
        h = C.c_void_p()
        rc = self.L.rt_server_create(ctx.h if ctx else None, C.byref(cfg), C.byref(h))
        if rc != RT_OK:
            raise RtError(rc, "rt_server_create")
        self.h = h

    def close(self):
        if self.h:
            self.L.rt_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int, what: str):
        if rc != RT_OK:
            raise RtError(rc, f"{what}: {self.L.rt_server_last_error(self.h).decode()}")

    def add_scene(self, key: int, xyz: np.ndarray, idx: np.ndarray, depth: int):
        xyz = np.ascontiguousarray(xyz, np.float32)
        idx = np.ascontiguousarray(idx, np.uint32)
        self.check(self.L.rt_server_add_scene(self.h, key, _ptr(xyz), xyz.size // 3, _ptr(idx),
                                              idx.size // 3, depth), "rt_server_add_scene")

    def submit(self, scene_key, view, cam, width, height, mode=MODE_GI, spp=1, max_bounces=4,
               albedo=0.7, sky=1.0, ambient=0.1, lights=None, seed=1, arrival=0.0) -> int:
        r = Request()
        r.scene_key, r.view, r.cam, r.width, r.height = scene_key, view, cam, width, height
        r.mode, r.spp, r.max_bounces = mode, spp, max_bounces
        r.albedo, r.sky, r.ambient = albedo, sky, ambient
        arr = np.ascontiguousarray(np.asarray(lights if lights is not None else [], np.float32).reshape(-1))
        r.lights = arr.ctypes.data_as(C.POINTER(C.c_float)) if arr.size else None
        r.nlights = arr.size // 3
        r.seed, r.arrival = seed, arrival
        rid = C.c_uint64()
        self.check(self.L.rt_server_submit(self.h, C.byref(r), C.byref(rid)), "rt_server_submit")
        return rid.value

    def step(self, now: float) -> float:
        """One scheduling step at `now`; returns its duration in seconds."""
        dt = C.c_double()
        self.check(self.L.rt_server_step(self.h, now, C.byref(dt)), "rt_server_step")
        return dt.value

    def poll(self, cap: int = 1024) -> list:
        out = (RequestResult * cap)()
        n = C.c_uint32()
        self.check(self.L.rt_server_poll(self.h, out, cap, C.byref(n)), "rt_server_poll")
        return [out[k].as_dict() for k in range(n.value)]

    def status(self) -> dict:
        st = ServerStats()
        self.check(self.L.rt_server_status(self.h, C.byref(st)), "rt_server_status")
        return st.as_dict()

    def read_stream(self, scene_key, view, width, height, mode):
        out = np.empty((height, width, 4), np.float32)
        n = C.c_uint32()
        self.check(self.L.rt_server_read_stream(self.h, scene_key, view, width, height, mode, _ptr(out),
                                                C.byref(n)), "rt_server_read_stream")
        return out, n.value

    def run(self, arrivals, now: float = 0.0, max_steps: int = 10_000_000, progress=None) -> tuple:
        """Steps on a simulated clock until every submitted request is
        finished: each step advances the clock by its duration, an idle server
        jumps to the next arrival.  arrivals: the submitted arrival times.
        progress (optional) is called with (clock, status) about every 10 s of
        wall time.  Returns (results, end time)."""
        import time as _time

        pending_t = sorted(arrivals)
        results, t, k = [], now, 0
        next_report = _time.perf_counter() + 10.0
        for _ in range(max_steps):
            st = self.status()
            if progress is not None and _time.perf_counter() >= next_report:
                progress(t, st)
                next_report += 10.0
            if st["pending"] == 0 and st["inflight"] == 0:
                break
            dt = self.step(t)
            results += self.poll()
            if dt > 0.0:
                t += dt
                continue
            st = self.status()
            if st["inflight"] == 0:
                while k < len(pending_t) and pending_t[k] <= t:
                    k += 1
                if st["wake"] > t:  # arrived requests held for the batch window
                    t = min(st["wake"], pending_t[k]) if k < len(pending_t) else st["wake"]
                    continue
                if k == len(pending_t):
                    break  # nothing left that could arrive (rejected or waiting for memory)
                t = pending_t[k]
        results += self.poll()
        return results, t
