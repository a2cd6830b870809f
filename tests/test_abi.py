"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/rtcore.h, and its host-only parts (scene generators,
camera helper, argument errors) behave."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "rtcore.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("rt_create", "rt_lbvh_build", "rt_lbvh_read", "rt_trace", "rt_render_direct",
                 "rt_scene_upload", "rt_scene_transform", "rt_gen_torus"):
        assert must in names


def test_library_exports_every_declared_symbol(rtcore):
    lib = rtcore.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", rtcore.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (rt_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported


def test_library_has_gfx950_code_object(rtcore):
    # the offload bundle of our own build names its target triple
    blob = open(rtcore.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_torus_generator(rtcore):
    xyz, idx = rtcore.gen_torus()
    assert xyz.shape == (128 * 64, 3) and idx.shape == (16384, 3)
    assert idx.max() < len(xyz)
    r_major = np.hypot(xyz[:, 0], xyz[:, 2])
    tube = np.hypot(r_major - 1.0, xyz[:, 1])
    assert np.all(np.abs(tube - 0.3) < 0.003)
    xyz2, idx2 = rtcore.gen_torus()
    assert np.array_equal(xyz, xyz2) and np.array_equal(idx, idx2)
    xyz3, _ = rtcore.gen_torus(seed=2)
    assert not np.array_equal(xyz, xyz3)


def test_triangle_soup_generator(rtcore):
    xyz, idx = rtcore.gen_triangle_soup(262144, 0.004, seed=2)
    assert idx.shape == (262144, 3) and xyz.shape == (3 * 262144, 3)
    c = xyz.reshape(-1, 3, 3).mean(1)
    assert np.all(c > -0.05) and np.all(c < 1.05)
    e1 = xyz[1::3] - xyz[0::3]
    assert abs(float(e1.std()) - 0.004) < 2e-4


def test_sphere_and_mix_generators(rtcore):
    xyz, idx = rtcore.gen_noisy_sphere(65, 64, 0.05, seed=3)
    assert len(idx) == 2 * 64 * 64
    r = np.linalg.norm(xyz, axis=1)
    assert r.min() > 0.95 - 1e-6 and r.max() < 1.05 + 1e-6
    xyz, idx = rtcore.gen_object_mix(5000, seed=4)
    assert len(idx) == 5000 and idx.max() < len(xyz)


def test_camera_look_at_basis(rtcore):
    cam = rtcore.camera_look_at((0, 0.5, 3), (0, 0, 0), (0, 1, 0), 60.0, 16 / 9)
    c = cam.as_array()
    f, r, u = c[3:6], c[6:9], c[9:12]
    for v in (f, r, u):
        assert abs(np.linalg.norm(v) - 1) < 1e-6
    assert abs(f @ r) < 1e-6 and abs(f @ u) < 1e-6 and abs(r @ u) < 1e-6
    assert abs(c[13] - np.tan(np.deg2rad(30))) < 1e-6 and abs(c[12] - c[13] * 16 / 9) < 1e-6


def test_create_without_gpu_reports_status(rtcore):
    h = ctypes.c_void_p()
    rc = rtcore.lib().rt_create(ctypes.byref(h), 0)
    if rc == rtcore.RT_OK:
        rtcore.lib().rt_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == rtcore.RT_ERR_NODEV and not h.value
    assert rtcore.lib().rt_last_error(None) == b"null context"


def test_python_constants_match_the_header(rtcore):
    """Every RT_* enum value of include/rtcore.h that rtcore.py mirrors (the
    name without the RT_ prefix) has the header's value; the flags and
    switches the tests and facades use are all mirrored."""
    import os
    import re

    src = open(os.path.join(os.path.dirname(__file__), "..", "include", "rtcore.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    values = {}
    for body in re.findall(r"enum\s*\w*\s*\{(.*?)\}", src, flags=re.S):
        for name, expr in re.findall(r"(RT_[A-Z0-9_]+)\s*=\s*([^,]+)", body):
            values[name] = eval(re.sub(r"(\d+)u\b", r"\1", expr.strip()), {}, {})
    assert len(values) > 20
    # status codes keep the prefix in Python (rtcore.RT_OK), the rest drop it
    py = {n: (n if n.startswith(("RT_OK", "RT_ERR")) else n[3:]) for n in values}
    mirrored = {n: v for n, v in values.items() if hasattr(rtcore, py[n])}
    for n, v in mirrored.items():
        assert getattr(rtcore, py[n]) == v, n
    for n in ("RT_FLAG_ANYHIT", "RT_FLAG_COUNT", "RT_FLAG_KEEP_HITS", "RT_FLAG_PERSISTENT",
              "RT_FLAG_START_AT_IROOT", "RT_WHITTED_ONE_PASS", "RT_WHITTED_SPLIT", "RT_WHITTED_QUADS", "RT_BVH_SAH",
              "RT_STRUCTURE_LBVH", "RT_MODE_GI", "RT_OK", "RT_ERR_BUDGET"):
        assert n in mirrored, n


STRUCTS = {"BuildTimes": "rt_build_times", "Profile": "rt_profile", "Camera": "rt_camera",
           "RenderDesc": "rt_render_desc", "WhittedDesc": "rt_whitted_desc", "Model": "rt_model",
           "BvhInfo": "rt_bvh_info", "Tuning": "rt_tuning", "JobDesc": "rt_job_desc",
           "ServerConfig": "rt_server_config", "Request": "rt_request", "RequestResult": "rt_request_result",
           "ServerStats": "rt_server_stats", "FrameStats": "rt_frame_stats"}
DTYPES = {"NODE_DTYPE": "rt_node32", "RAY_DTYPE": "rt_ray48", "HIT_DTYPE": "rt_hit24",
          "MATERIAL_DTYPE": "rt_material"}


def test_python_structures_match_the_header(rtcore, tmp_path):
    """Every ctypes structure and record dtype of the Python facade has the
    size and field offsets of its C struct in include/rtcore.h (compiled
    with the host C compiler), so a field added to one side alone fails."""
    want, lines = {}, []
    for py, c in STRUCTS.items():
        cls = getattr(rtcore, py)
        want[f"{c} sizeof"] = ctypes.sizeof(cls)
        lines.append(f'printf("{c} sizeof %zu\\n", sizeof({c}));')
        for name, _ in cls._fields_:
            want[f"{c} {name}"] = getattr(cls, name).offset
            lines.append(f'printf("{c} {name} %zu\\n", offsetof({c}, {name}));')
    for py, c in DTYPES.items():
        dt = getattr(rtcore, py)
        want[f"{c} sizeof"] = dt.itemsize
        lines.append(f'printf("{c} sizeof %zu\\n", sizeof({c}));')
        for name in dt.names:
            want[f"{c} {name}"] = dt.fields[name][1]
            lines.append(f'printf("{c} {name} %zu\\n", offsetof({c}, {name}));')
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include "rtcore.h"\nint main(void) {\n' + "\n".join(lines) +
                   "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        key, v = line.rsplit(" ", 1)
        got[key] = int(v)
    assert got == want
