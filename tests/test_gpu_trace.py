"""GPU traversal parity against the oracle's scalar transcription of
LBVH_IntersectP + Intersection.hlsl, and the direct-lighting frame against the
oracle's reference-semantics shading (shadow rays closest-hit from the hit
leaf).  Hits (id, u, v, t, visit counter, leaf) are expected bit-exact: the
device computes the same IEEE single operations in the same order and visits
nodes in the same order.  The acceptance bar of the path is written beside
each check: >= 99.99 % hit ids with t within 1e-5 relative otherwise, and
PSNR >= 50 dB for the direct-lighting image."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T_REL = 1e-5


def psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10 * np.log10(1.0 / mse)


def compare_hits(g, r, tag):
    same = (g["tri"] == r["tri"])
    frac = same.mean()
    print(f"{tag}: rays={len(g)} id-match={frac:.6f} exact={np.mean(g.tobytes() == r.tobytes())}",
          flush=True)
    assert frac >= 0.9999, tag
    # mismatches only at ties of t
    bad = ~same
    if bad.any():
        assert np.all(np.abs(g["t"][bad] - r["t"][bad]) <= T_REL * np.abs(r["t"][bad]))
    hit = same & (r["tri"] >= 0)
    assert np.all(np.abs(g["t"][hit] - r["t"][hit]) <= T_REL * np.abs(r["t"][hit]))
    assert np.all(np.abs(g["u"][hit] - r["u"][hit]) <= T_REL * np.maximum(np.abs(r["u"][hit]), 1e-1))
    assert np.all(np.abs(g["v"][hit] - r["v"][hit]) <= T_REL * np.maximum(np.abs(r["v"][hit]), 1e-1))


@pytest.fixture(scope="module")
def torus(ctx, oracle, rtcore):
    xyz, idx = rtcore.gen_torus()
    s = ctx.add_scene(xyz, idx, 20)
    ctx.build(s)
    ref = oracle.build_scene(xyz, idx, 20)
    yield s, xyz, idx, ref
    ctx.scene_destroy(s)


@pytest.fixture(scope="module")
def soup262k(ctx, oracle, rtcore):
    xyz, idx = rtcore.gen_triangle_soup(262144, 0.004, 2)
    s = ctx.add_scene(xyz, idx, 20)
    ctx.build(s)
    ref = oracle.build_scene(xyz, idx, 20)
    yield s, xyz, idx, ref
    ctx.scene_destroy(s)


def test_trace_torus_primary_512(ctx, oracle, rtcore, torus):
    s, xyz, idx, ref = torus
    cam = rtcore.camera_look_at((0, 0.5, 3), (0, 0, 0), (0, 1, 0), 60.0, 1.0)
    rays = oracle.raygen_pinhole(cam.as_array(), 512, 512)
    g = ctx.trace(s, rays)
    r = oracle.trace(ref["nodes"], ref["ids"], xyz, idx, rays, nthreads=8)
    compare_hits(g, r, "torus primary")
    assert np.array_equal(g, r)  # bit-exact, visit counters and leaves included


def test_trace_soup_random_rays(ctx, oracle, soup262k):
    s, xyz, idx, ref = soup262k
    rng = np.random.default_rng(21)
    import rtcore

    n = 200000
    rays = np.zeros(n, rtcore.RAY_DTYPE)
    rays["o"] = rng.random((n, 3)).astype(np.float32) * 1.4 - 0.2
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = 1e16
    rays["tri"] = -1
    rays["tri"][::97] = rng.integers(0, len(idx), size=len(rays["tri"][::97]))  # some self ids
    g = ctx.trace(s, rays)
    r = oracle.trace(ref["nodes"], ref["ids"], xyz, idx, rays, nthreads=8)
    compare_hits(g, r, "soup random")
    assert np.array_equal(g, r)


def test_trace_anyhit_and_dead_rays(ctx, oracle, rtcore, soup262k):
    s, xyz, idx, ref = soup262k
    cam = rtcore.camera_look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 60.0, 1.0)
    rays = oracle.raygen_pinhole(cam.as_array(), 256, 256)
    rays["tmax"] = 2.0
    rays["tri"][::5] = -2  # dead rays are skipped (CSIntersections.hlsl:23,37)
    g = ctx.trace(s, rays, flags=rtcore.FLAG_ANYHIT)
    live = rays["tri"] > -2
    r = oracle.trace(ref["nodes"], ref["ids"], xyz, idx, rays[live], anyhit=True, nthreads=8)
    assert np.all(g["tri"][~live] == -2)
    gl = g[live]
    for f in ("tri", "u", "v", "t", "root"):
        assert np.array_equal(gl[f], r[f]), f
    # any-hit stops early; the device tests both child boxes when it enters a
    # node, so its box counter can include siblings the trail loop never
    # reaches.  Triangle tests are the same sequence.
    assert np.array_equal(gl["visited"] >> 16, r["visited"] >> 16)
    assert np.all((gl["visited"] & 0xFFFF) >= (r["visited"] & 0xFFFF))


def test_trace_counts_match_visit_counters(ctx, oracle, rtcore, torus):
    s, xyz, idx, ref = torus
    cam = rtcore.camera_look_at((0, 0.5, 3), (0, 0, 0), (0, 1, 0), 60.0, 1.0)
    rays = oracle.raygen_pinhole(cam.as_array(), 128, 128)
    g, st = ctx.trace(s, rays, flags=rtcore.FLAG_COUNT, want_stats=True)
    assert st["box_tests"] == int((g["visited"] & 0xFFFF).sum())
    assert st["tri_tests"] == int((g["visited"] >> 16).sum())


@pytest.mark.parametrize("scene,cam_pos,target,light", [
    ("torus", (0, 0.5, 3), (0, 0, 0), [(0, 0, 0)]),
    ("soup", (0.5, 0.5, -1.5), (0.5, 0.5, 0.5), [(0.5, 2.0, -1.0), (1.5, 0.5, 0.5)]),
])
def test_render_direct_parity(ctx, oracle, rtcore, torus, soup262k, scene, cam_pos, target, light):
    s, xyz, idx, ref = torus if scene == "torus" else soup262k
    W, H = (512, 512) if scene == "torus" else (480, 270)
    cam = rtcore.camera_look_at(cam_pos, target, (0, 1, 0), 60.0, W / H)
    st = ctx.render_direct(s, cam, W, H, light, flags=rtcore.FLAG_KEEP_HITS)
    rgba, hits, occ = ctx.frame_read(W, H, len(light), hits=True)
    rays = oracle.raygen_pinhole(cam.as_array(), W, H)
    r = oracle.trace(ref["nodes"], ref["ids"], xyz, idx, rays, nthreads=8)
    compare_hits(hits, r, f"{scene} frame primary")
    assert np.array_equal(hits, r)
    rrgba, rocc = oracle.render_direct(ref["nodes"], ref["ids"], xyz, idx, rays, r, light, nthreads=8)
    occ_match = (occ[:, :len(light)] == rocc[:, :len(light)]).mean()
    p = psnr(rgba[:, :3], rrgba[:, :3])
    print(f"{scene}: shadow rays={st['shadow_rays']} occlusion match={occ_match:.6f} PSNR={p:.2f} dB",
          flush=True)
    assert occ_match >= 0.9999
    assert p >= 50.0
    assert st["shadow_rays"] == int((rocc[:, :len(light)] != 2).sum())


def test_facade_demo_binary(rtcore, tmp_path):
    """The C++ facades end to end: direct frames with rebuilds, a Whitted
    frame of 2 bounces on the SAH BVH, a PNG screenshot, one IntersectP."""
    import json
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(rtcore.LIB_PATH), "rt_frame_demo")
    shot = tmp_path / "shot.png"
    raw = tmp_path / "world.f32"
    out = subprocess.run([exe, "3", str(shot), str(raw)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    assert rec["samples"] == 5 and rec["keyboard_id"] == 9 and rec["name"] == "LBVH"
    assert rec["probe_tri"] >= 0 and 2.0 < rec["probe_t"] < 3.0
    assert rec["rays"] >= 3 * 512 * 512
    assert rec["whitted_rays"] > 512 * 512
    assert shot.read_bytes()[:8] == b"\x89PNG\r\n\x1a\n"
    # the frame of RayTracerHip::SetCamera(world): the reference's camera-matrix
    # rays (CSPrimaryRays.hlsl:39-52) against the oracle's, light at the origin
    import oracle

    world = np.array(rec["world"], np.float32)
    got = np.fromfile(raw, np.float32).reshape(-1, 4)
    xyz, idx = rtcore.gen_torus(seed=7)
    ref = oracle.build_scene(xyz, idx, 20)
    rays = oracle.raygen_world(world, 512, 512)
    hits = oracle.trace(ref["nodes"], ref["ids"], xyz, idx, rays, nthreads=8)
    want, _ = oracle.render_direct(ref["nodes"], ref["ids"], xyz, idx, rays, hits, [(0, 0, 0)], nthreads=8)
    same = float(np.mean(np.all(got == want, axis=1)))
    print(f"facade world camera: hits {(hits['tri'] >= 0).mean():.3f}, identical pixels {same:.6f}, "
          f"PSNR {psnr(got[:, :3], want[:, :3]):.2f} dB", flush=True)
    assert (hits["tri"] >= 0).mean() > 0.05
    assert same >= 0.9999 and psnr(got[:, :3], want[:, :3]) >= 50.0


@pytest.mark.parametrize("form", [1, 2])
def test_render_direct_forms_equal(ctx, rtcore, soup262k, form):
    """Every traversal form gives the same frame, hits, occlusion and counters as the loop form."""
    s = soup262k[0]
    W, H = 333, 177
    cam = rtcore.camera_look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 60.0, W / H)
    lights = [(0.5, 2.0, -1.0), (1.5, 0.5, 0.5), (-0.5, 0.5, 0.2)]
    saved = ctx.tuning()
    out = []
    try:
        for f in (0, form):
            ctx.set_tuning(128, 12, 4, f)
            st = ctx.render_direct(s, cam, W, H, lights, flags=rtcore.FLAG_KEEP_HITS | rtcore.FLAG_COUNT)
            out.append((st, ctx.frame_read(W, H, len(lights), hits=True)))
    finally:
        ctx.set_tuning(*saved)
    (st0, a), (st1, b) = out
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    for k in ("primary_rays", "shadow_rays", "box_tests", "tri_tests"):
        assert st0[k] == st1[k], k


@pytest.mark.parametrize("view", ["sparse", "dense"])
def test_render_direct_shadow_groups_equal(ctx, rtcore, soup262k, view):
    """Uncounted frames of the shared-fetch form trace each wave's shadow rays
    on groups of lanes (up to 16 per ray, each searching every k-th subtree
    below heap depth 8, stopping when one finds an occluder).  Frame, kept
    hits and occlusion bytes equal the loop form's one lane per ray, in a view
    with few shadow rays per wave (large groups) and in the 1080p soup frame
    (about half the lanes of a wave cast one: groups of one or two)."""
    s = soup262k[0]
    if view == "sparse":
        W, H = 401, 233
        cam = rtcore.camera_look_at((1.8, 0.9, -0.6), (0.5, 0.5, 0.5), (0, 1, 0), 50.0, W / H)
        lights = [(0.5, 2.0, -1.0), (1.5, 0.5, 0.5), (-0.5, 0.5, 0.2)]
    else:
        W, H = 1920, 1080
        cam = rtcore.camera_look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 60.0, W / H)
        lights = [(0.5, 2.5, -1.0)]
    saved = ctx.tuning()
    out = []
    try:
        for form, block in ((0, 128), (2, 128), (2, 64)):
            ctx.set_tuning(block, 20, 4, form)
            for flags in (0, rtcore.FLAG_KEEP_HITS):
                st = ctx.render_direct(s, cam, W, H, lights, flags=flags)
                out.append((form, block, flags, st["shadow_rays"],
                            ctx.frame_read(W, H, len(lights), hits=bool(flags))))
    finally:
        ctx.set_tuning(*saved)
    ref = {flags: (n, arrs) for form, block, flags, n, arrs in out if form == 0}
    assert ref[0][0] > 0
    for form, block, flags, n, arrs in out:
        assert n == ref[flags][0], (form, block, flags)
        for x, y in zip(ref[flags][1], arrs):
            assert (x is None and y is None) or np.array_equal(x, y), (form, block, flags)


@pytest.mark.parametrize("block", [64, 128, 256])
def test_render_direct_tile_order_keeps_the_frame(ctx, rtcore, soup262k, block):
    """rt_render_direct dispatches the tiles of a repeated frame heaviest
    first (from the last frame's per-tile costs): the first frame of a shape
    runs in block order, the next ones in cost order, and every frame is the
    same bit for bit, counters included.  A camera move, a size change and a
    shape change keep it so."""
    s = soup262k[0]
    W, H = 401, 233
    cams = [rtcore.camera_look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 60.0, W / H),
            rtcore.camera_look_at((1.8, 0.9, -0.6), (0.5, 0.5, 0.5), (0, 1, 0), 50.0, W / H)]
    lights = [(0.5, 2.0, -1.0), (1.5, 0.5, 0.5)]
    saved = ctx.tuning()
    try:
        ctx.set_tuning(block, 26 if block == 128 else 12, 4, 2)
        for cam in cams:
            ref = None
            for rep in range(3):
                st = ctx.render_direct(s, cam, W, H, lights, flags=rtcore.FLAG_KEEP_HITS | rtcore.FLAG_COUNT)
                out = ctx.frame_read(W, H, len(lights), hits=True)
                if ref is None:
                    ref = (st, out)
                    continue
                for x, y in zip(ref[1], out):
                    assert np.array_equal(x, y), (block, rep)
                for k in ("primary_rays", "shadow_rays", "box_tests", "tri_tests"):
                    assert ref[0][k] == st[k], (block, rep, k)
        # a smaller frame in the same buffers' place, then the first size again
        ctx.render_direct(s, cams[0], 97, 55, lights)
        ctx.render_direct(s, cams[0], W, H, lights, flags=rtcore.FLAG_KEEP_HITS)
        again = ctx.frame_read(W, H, len(lights), hits=True)
        ctx.render_direct(s, cams[0], W, H, lights, flags=rtcore.FLAG_KEEP_HITS)
        twice = ctx.frame_read(W, H, len(lights), hits=True)
        for x, y in zip(again, twice):
            assert np.array_equal(x, y)
    finally:
        ctx.set_tuning(*saved)


def quad_frames(ctx, rtcore, s, cams, W, H, lights, block, form, quads):
    """rgba and shadow-ray counts of each camera's frame with one lane per
    pixel and with quads on the `quads` heaviest tiles; every frame is drawn
    three times (the first in block order, the others in cost order, which
    is when quads run)."""
    lds = 26 if block == 128 else 12
    out = []
    for q in (0, quads):
        ctx.set_tuning(block, lds, 4, form, q)
        frames = []
        for cam in cams:
            for rep in range(3):
                st = ctx.render_direct(s, cam, W, H, lights)
                frames.append((ctx.frame_read(W, H, len(lights))[0], st["shadow_rays"], st["primary_rays"]))
        out.append(frames)
    return out


@pytest.mark.parametrize("block,form", [(64, 2), (64, 0), (128, 1), (256, 2)])
def test_render_direct_quads_keep_the_frame(ctx, rtcore, soup262k, torus, block, form):
    """Quads (four lanes per pixel on the heaviest tiles, each lane searching
    every fourth subtree of heap depth 8 along the ray) give the frame of one
    lane per pixel bit for bit: on the soup and on the torus mesh (whose rays
    meet shared edges, where the tie rule decides), in a ragged image, with
    one, a few or all tiles in quads, and the same shadow-ray count."""
    W, H = 401, 233
    cams = [rtcore.camera_look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 60.0, W / H),
            rtcore.camera_look_at((1.8, 0.9, -0.6), (0.5, 0.5, 0.5), (0, 1, 0), 50.0, W / H)]
    tcams = [rtcore.camera_look_at((0.0, 0.5, 3.0), (0, 0, 0), (0, 1, 0), 60.0, W / H),
             rtcore.camera_look_at((2.0, 1.5, 0.5), (0, 0, 0), (0, 1, 0), 45.0, W / H)]
    lights = [(0.5, 2.0, -1.0), (1.5, 0.5, 0.5)]
    saved = ctx.tuning()
    try:
        for scene, cs in ((soup262k[0], cams), (torus[0], tcams)):
            for quads in (1, 7, 1 << 20):
                one, quad = quad_frames(ctx, rtcore, scene, cs, W, H, lights, block, form, quads)
                for k, ((a, sa, pa), (b, sb, pb)) in enumerate(zip(one, quad)):
                    assert np.array_equal(a, b), (quads, k, int(np.any(a != b, axis=1).sum()))
                    assert (sa, pa) == (sb, pb), (quads, k)
    finally:
        ctx.set_tuning(*saved)


def test_render_direct_quads_small_and_shallow_trees(ctx, rtcore):
    """Trees whose leaves sit above heap depth 8 (one triangle: the root is a
    leaf; a handful: every leaf is an item of its own; a chain as deep as the
    heap allows, with a leaf at every level) and an SAH BVH scene (which
    keeps one lane per pixel): quads give the one-lane frame."""
    W, H = 160, 90
    lights = [(0.3, 3.0, -2.0)]
    saved = ctx.tuning()
    try:
        scenes = []
        for n in (1, 2, 5, 40):
            xyz, idx = rtcore.gen_triangle_soup(n, 0.2, 7 + n)
            scenes.append((ctx.add_scene(xyz, idx, 20), False))
        cx, ci = chain_scene()
        scenes.append((ctx.add_scene(cx, ci, 26), False))
        xyz, idx = rtcore.gen_triangle_soup(3000, 0.01, 3)
        scenes.append((ctx.add_scene(xyz, idx, 12), True))
        cams = [rtcore.camera_look_at((0.5, 0.5, -1.2), (0.5, 0.5, 0.5), (0, 1, 0), 70.0, W / H),
                rtcore.camera_look_at((-0.01, -0.01, -0.01), (0.01, 0.01, 0.01), (0, 1, 0), 60.0, W / H)]
        for s, sah in scenes:
            if sah:
                ctx.bvh_build(s)
            else:
                ctx.build(s)
            for block, form in ((64, 2), (128, 0)):
                one, quad = quad_frames(ctx, rtcore, s, cams, W, H, lights, block, form, 1 << 20)
                for k, ((a, sa, pa), (b, sb, pb)) in enumerate(zip(one, quad)):
                    assert np.array_equal(a, b), (s, block, form, k)
                    assert (sa, pa) == (sb, pb)
        for s, _ in scenes:
            ctx.scene_destroy(s)
    finally:
        ctx.set_tuning(*saved)


def test_tuning_quad_tiles_is_checked(ctx, rtcore):
    """rt_tuning_set takes quad_tiles in 0..2^24 and keeps it; other values
    are rejected and leave the tuning as it was."""
    saved = ctx.tuning()
    try:
        ctx.set_tuning(64, 20, 4, 2, 256)
        assert ctx.tuning() == (64, 20, 4, 2, 256)
        for bad in (-1, (1 << 24) + 1):
            with pytest.raises(rtcore.RtError):
                ctx.set_tuning(64, 20, 4, 2, bad)
            assert ctx.tuning() == (64, 20, 4, 2, 256)
    finally:
        ctx.set_tuning(*saved)


def test_render_direct_views_keep_their_orders(ctx, rtcore, soup262k):
    """Each view (scene, camera, shape) keeps its own tile order; ten cameras
    taken in turn, more than the eight order slots, with quads on their
    heaviest tiles: every frame equals the camera's first frame (block
    order, one lane per pixel)."""
    import math

    s = soup262k[0]
    W, H = 200, 120
    lights = [(0.5, 2.0, -1.0)]
    cams = [rtcore.camera_look_at((0.5 + 1.8 * math.cos(a), 0.7, 0.5 + 1.8 * math.sin(a)), (0.5, 0.5, 0.5),
                                  (0, 1, 0), 55.0, W / H) for a in np.linspace(0.0, 6.0, 10)]
    saved = ctx.tuning()
    try:
        ctx.set_tuning(64, 20, 4, 2, 0)
        ref = []
        for c in cams:
            ctx.render_direct(s, c, W, H, lights)
            ref.append(ctx.frame_read(W, H, 1)[0])
        ctx.set_tuning(64, 20, 4, 2, 64)
        for rnd in range(3):
            for k in (list(range(10)) if rnd != 1 else [0, 1, 0, 1, 2, 9, 2]):
                ctx.render_direct(s, cams[k], W, H, lights)
                assert np.array_equal(ctx.frame_read(W, H, 1)[0], ref[k]), (rnd, k)
    finally:
        ctx.set_tuning(*saved)


def chain_scene():
    """Triangles whose Morton codes are 2^29, 2^28, ..., 1, 0: each split
    peels one off, so the tree is a chain as deep as the heap allows."""
    tris = []
    for b in range(29, -1, -1):
        axis, j = 2 - (b % 3), b // 3
        c = np.zeros(3)
        c[axis] = (2 ** j + 0.5) / 1024.0
        tris.append(c)
    tris.append(np.full(3, 0.2 / 1024.0))
    tris.append(np.full(3, 1.0 - 0.5 / 1024.0))  # far corner: fixes the scene box
    tris = np.array(tris)
    e = 0.2 / 1024.0
    xyz = np.concatenate([tris + [0, 0, 0], tris + [e, 0, 0], tris + [0, e, 0]], axis=1).reshape(-1, 3)
    xyz = np.vstack([xyz, [[0, 0, 0], [1e-4, 0, 0], [0, 1e-4, 0]]]).astype(np.float32)  # near corner
    idx = np.arange(len(xyz), dtype=np.uint32).reshape(-1, 3)
    return xyz, idx


def test_deep_heap_under_every_shape(ctx, oracle, rtcore):
    """A chain 25 levels deep (heap depth 26): launch shapes whose LDS stack
    is shorter than depth - 1 either spill (16 or fewer entries) or are
    replaced by one that fits (20 and 22 by 26); every shape and form gives
    the oracle's hits on rays that walk down the chain."""
    xyz, idx = chain_scene()
    s = ctx.add_scene(xyz, idx, 26)
    ctx.build(s)
    ref = oracle.build_scene(xyz, idx, 26)
    assert int(np.floor(np.log2(oracle.reachable(ref["nodes"]))).max()) == 25
    rng = np.random.default_rng(5)
    n = 4096
    rays = np.zeros(n, oracle.RAY_DTYPE)
    o = np.array([-0.01, -0.01, -0.01], np.float32)
    tgt = rng.uniform(0.0, 0.02, (n, 3)).astype(np.float32)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["o"] = o
    rays["d"] = d
    rays["tmax"] = 1e16
    rays["tri"] = -1
    r = oracle.trace(ref["nodes"], ref["ids"], xyz, idx, rays)
    assert (r["visited"] & 0xFFFF).max() >= 40  # the rays walk far down the chain
    saved = ctx.tuning()
    try:
        for form in (0, 1, 2):
            for block, lds in [(64, 16), (64, 20), (128, 12), (128, 16), (128, 20), (128, 22), (128, 26),
                               (256, 12)]:
                ctx.set_tuning(block, lds, 4, form)
                assert np.array_equal(ctx.trace(s, rays), r), (block, lds, form)
    finally:
        ctx.set_tuning(*saved)
    ctx.scene_destroy(s)


def test_trace_from_iroot_matches_oracle(ctx, oracle, rtcore):
    """RT_FLAG_START_AT_IROOT: each ray starts at the heap node in its hit
    record (the previous intersection's iRoot), as CSComputeIntersections
    starts LBVH_IntersectP.  Bounce-like and shadow rays from the hit points,
    started at the hit leaves, equal the oracle's trail loop from the same
    nodes: ids, u, v, t, the visit counter and the new iRoot, closest and any
    hit, under a spilling and a non-spilling launch shape.  Starts 0 and 1 are
    the root; a start below a leaf starts at that leaf.  Any hit: the same
    hits, and box counts at least the trail loop's, as from the root."""
    xyz, idx = rtcore.gen_triangle_soup(60000, 0.008, seed=12)
    s = ctx.add_scene(xyz, idx, 18)
    ctx.build(s)
    ref = oracle.build_scene(xyz, idx, 18)
    nodes = ref["nodes"]
    cam = rtcore.camera_look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 60.0, 1.0)
    prim = oracle.raygen_pinhole(cam.as_array(), 128, 128)
    h0 = ctx.trace(s, prim)
    hit = h0["tri"] >= 0
    assert hit.mean() > 0.2 and (h0["root"][hit] > 1).all()
    p = prim["o"][hit] + h0["t"][hit][:, None] * prim["d"][hit]
    rng = np.random.default_rng(5)
    d = rng.normal(size=(len(p), 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    bounce = np.zeros(len(p), oracle.RAY_DTYPE)
    bounce["o"], bounce["d"], bounce["refl"], bounce["tmax"], bounce["tri"] = p, d, 1.0, 1e16, h0["tri"][hit]
    L = np.array([0.5, 2.0, -1.0], np.float32)
    ld = (L - p).astype(np.float32)
    dist = np.linalg.norm(ld, axis=1).astype(np.float32)
    shadow = bounce.copy()
    shadow["d"] = (ld / dist[:, None]).astype(np.float32)
    shadow["tmax"] = dist
    start = h0["root"][hit]
    saved = ctx.tuning()
    try:
        for shape in [(128, 20, 4, 2), (64, 16, 4, 0)]:
            ctx.set_tuning(*shape)
            for rays, anyhit in ((bounce, False), (shadow, False), (shadow, True)):
                fl = rtcore.FLAG_ANYHIT if anyhit else 0
                g = ctx.trace(s, rays, flags=fl, start=start)
                r = oracle.trace(nodes, ref["ids"], xyz, idx, rays, start=start, anyhit=anyhit, nthreads=8)
                if not anyhit:
                    assert np.array_equal(g, r), (shape, anyhit)
                else:  # an early stop: boxes the device tested up front may be more (as from the root)
                    for f in ("tri", "u", "v", "t", "root"):
                        assert np.array_equal(g[f], r[f]), (shape, f)
                    assert np.array_equal(g["visited"] >> 16, r["visited"] >> 16)
                    assert np.all((g["visited"] & 0xFFFF) >= (r["visited"] & 0xFFFF))
                from_root = oracle.trace(nodes, ref["ids"], xyz, idx, rays, anyhit=anyhit, nthreads=8)
                assert not np.array_equal(g["visited"], from_root["visited"])  # the start matters
        # 0 and 1 are the root
        plain = ctx.trace(s, bounce)
        for root in (0, 1):
            assert np.array_equal(ctx.trace(s, bounce, start=np.full(len(bounce), root, np.int32)), plain)
        # below a leaf: that leaf (the first leaf on the path of the start's bits)
        leaf = (nodes["prim_count"] & 4) != 0
        odd = np.concatenate([2 * start, 4 * start + 3, np.full(8, (1 << 30) + 12345, np.int64)]).astype(np.int64)
        clamped = []
        for v in odd:
            lv = int(v).bit_length() - 1
            node = 1
            for k in range(lv - 1, -1, -1):
                if leaf[node]:
                    break
                node = 2 * node + ((int(v) >> k) & 1)
            clamped.append(node)
        many = np.concatenate([bounce, bounce, bounce[:8]])
        g = ctx.trace(s, many, start=odd.astype(np.int32))
        r = oracle.trace(nodes, ref["ids"], xyz, idx, many, start=np.array(clamped, np.int32), nthreads=8)
        assert np.array_equal(g, r)
    finally:
        ctx.set_tuning(*saved)
    bvh = ctx.add_scene(xyz[:3000], idx[:1000] % 3000, 12)
    ctx.bvh_build(bvh)
    with pytest.raises(rtcore.RtError):
        ctx.trace(bvh, bounce[:4], start=np.ones(4, np.int32))
    ctx.scene_destroy(bvh)
    ctx.scene_destroy(s)


@pytest.mark.parametrize("config", ["soup262k_1080p", "sphere2m_1080p"])
def test_render_direct_full_size_parity(ctx, oracle, rtcore, config):
    """The benchmark frames at their full size against the oracle: the
    262,144-triangle soup at 1920x1080 with its light, and one of the
    2,097,152-triangle sphere's orbit cameras at 1920x1080 with 4 lights
    (depth 23).  Primary hits bit-exact, >= 99.99 % occlusion agreement,
    PSNR >= 50 dB, shadow-ray count equal; every frame is rendered twice (the
    second in the heaviest-tiles-first order) and both are checked, then
    twice more with quads on the 1024 heaviest tiles (the same image)."""
    import math

    W, H = 1920, 1080
    if config == "soup262k_1080p":
        xyz, idx = rtcore.gen_triangle_soup(262144, 0.004, 2)
        depth = 20
        cam = rtcore.camera_look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 60.0, W / H)
        lights = [(0.5, 2.5, -1.0)]
    else:
        xyz, idx = rtcore.gen_noisy_sphere(1025, 1024, 0.05, 3)
        depth = 23
        k = 3
        cam = rtcore.camera_look_at((2.5 * math.cos(2 * math.pi * k / 8), 0.0, 2.5 * math.sin(2 * math.pi * k / 8)),
                                    (0, 0, 0), (0, 1, 0), 60.0, W / H)
        lights = [(3.0, 3.0, 3.0), (-3.0, 3.0, 3.0), (3.0, 3.0, -3.0), (-3.0, -3.0, -3.0)]
    s = ctx.add_scene(xyz, idx, depth)
    ctx.build(s)
    ref = oracle.build_scene(xyz, idx, depth)
    rays = oracle.raygen_pinhole(cam.as_array(), W, H)
    r = oracle.trace(ref["nodes"], ref["ids"], xyz, idx, rays, nthreads=16)
    rrgba, rocc = oracle.render_direct(ref["nodes"], ref["ids"], xyz, idx, rays, r, lights, nthreads=16)
    nl = len(lights)
    for rep in range(2):
        st = ctx.render_direct(s, cam, W, H, lights, flags=rtcore.FLAG_KEEP_HITS)
        rgba, hits, occ = ctx.frame_read(W, H, nl, hits=True)
        assert np.array_equal(hits, r), rep
        occ_match = (occ[:, :nl] == rocc[:, :nl]).mean()
        p = psnr(rgba[:, :3], rrgba[:, :3])
        print(f"{config} [{rep}]: hits {(r['tri'] >= 0).mean():.3f}, shadow rays {st['shadow_rays']}, "
              f"occlusion match {occ_match:.6f}, PSNR {p:.2f} dB", flush=True)
        assert occ_match >= 0.9999 and p >= 50.0
        assert st["shadow_rays"] == int((rocc[:, :nl] != 2).sum())
    # the same frame with quads on its 1024 heaviest tiles: the image above
    saved = ctx.tuning()
    try:
        ctx.set_tuning(*saved[:4], 1024)
        for rep in range(2):
            st = ctx.render_direct(s, cam, W, H, lights)
            assert np.array_equal(ctx.frame_read(W, H, nl)[0], rgba), rep
            assert st["shadow_rays"] == int((rocc[:, :nl] != 2).sum())
    finally:
        ctx.set_tuning(*saved)
    ctx.scene_destroy(s)
