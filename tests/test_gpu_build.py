"""GPU build parity: Morton codes, sorted (code, id) arrays, scene box, heap
topology and boxes from the HIP build against the CPU oracle, bit for bit.

The oracle is run in presorted mode (stable (code, id) order, the order the
onesweep sort produces); its topology and boxes equal the reference's
in-place-partition mode (tests/test_oracle.py pins that), whose per-leaf id
multisets are checked here too.  Boxes are compared as float values (exact
min/max; -0.0 == +0.0)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def soup(n, sigma=0.004, seed=2):
    import rtcore

    return rtcore.gen_triangle_soup(n, sigma, seed)


def dup_grid(n, cells=6, seed=1):
    rng = np.random.default_rng(seed)
    c = np.floor(rng.random((n, 3)) * cells) / cells
    xyz = np.concatenate([c, c + [0.01, 0, 0], c + [0, 0.01, 0]]).astype(np.float32)
    idx = np.stack([np.arange(n), np.arange(n) + n, np.arange(n) + 2 * n], 1).astype(np.uint32)
    return xyz, idx


def check_build(ctx, oracle, xyz, idx, depth, check_partition_mode=False):
    print(f"build parity: nt={len(idx)} depth={depth}", flush=True)
    s = ctx.add_scene(xyz, idx, depth)
    try:
        ctx.build(s)
        nodes, codes, ids, box = ctx.read(s)
    finally:
        ctx.scene_destroy(s)
    ref = oracle.build_scene(xyz, idx, depth, presorted=True)
    assert np.array_equal(box, ref["scene"])
    assert np.array_equal(codes, ref["codes"])
    assert np.array_equal(ids, ref["ids"])
    if len(idx) == 0:
        return
    reach = oracle.reachable(ref["nodes"])
    g, r = nodes[reach], ref["nodes"][reach]
    assert np.array_equal(g["prim_count"], r["prim_count"])
    assert np.array_equal(g["prim_pos"], r["prim_pos"])
    assert np.all(g["mn"] == r["mn"]) and np.all(g["mx"] == r["mx"])
    # children of leaves inside the heap carry the invalid marker
    leaves = reach[(r["prim_count"] & 4) != 0]
    inner = leaves[leaves < (1 << (depth - 1))]
    for c in (2 * inner, 2 * inner + 1):
        assert np.all(nodes["prim_count"][c] == oracle.INVALID)
        assert np.all(nodes["prim_pos"][c] == oracle.INVALID)
    if check_partition_mode:
        cpu = oracle.build_scene(xyz, idx, depth, presorted=False)
        assert np.array_equal(cpu["nodes"][reach]["prim_count"], g["prim_count"])
        for h in leaves:
            pos, cnt = int(r[reach == h]["prim_pos"][0]), int(r[reach == h]["prim_count"][0]) >> 3
            assert sorted(ids[pos:pos + cnt]) == sorted(cpu["ids"][pos:pos + cnt])


def test_build_torus_config1(ctx, oracle, rtcore):
    xyz, idx = rtcore.gen_torus()
    check_build(ctx, oracle, xyz, idx, 20, check_partition_mode=True)


def test_build_soup_262144(ctx, oracle):
    xyz, idx = soup(262144)
    check_build(ctx, oracle, xyz, idx, 20)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 1000, 1023, 1024, 1025, 4097, 65537])
def test_build_ragged_sizes(ctx, oracle, n):
    xyz, idx = soup(n, sigma=0.01, seed=n)
    check_build(ctx, oracle, xyz, idx, 20)


@pytest.mark.parametrize("depth", [2, 3, 6, 10, 14])
def test_build_depth_limit(ctx, oracle, depth):
    xyz, idx = soup(50000, seed=depth)
    check_build(ctx, oracle, xyz, idx, depth)


def test_build_duplicate_codes(ctx, oracle):
    xyz, idx = dup_grid(20000)
    check_build(ctx, oracle, xyz, idx, 16, check_partition_mode=True)


def test_build_all_identical_triangles(ctx, oracle):
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    xyz = np.tile(tri, (500, 1))
    idx = np.arange(1500, dtype=np.uint32).reshape(500, 3)
    check_build(ctx, oracle, xyz, idx, 8)


def test_build_indexed_mesh_unreferenced_vertices(ctx, oracle, rtcore):
    # vertices outside every triangle must not widen the scene box
    xyz, idx = rtcore.gen_torus(nu=32, nv=16)
    xyz = np.concatenate([xyz, [[50, 50, 50], [-50, -50, -50]]]).astype(np.float32)
    check_build(ctx, oracle, xyz, idx, 12)


def test_build_sphere_2m(ctx, oracle, rtcore):
    xyz, idx = rtcore.gen_noisy_sphere(1025, 1024, 0.05, seed=3)
    assert len(idx) == 2097152
    check_build(ctx, oracle, xyz, idx, 22)


@pytest.mark.parametrize("kind", ["soup", "dup_grid"])
def test_build_one_big_scene_top_levels_per_launch(ctx, oracle, rtcore, kind):
    """A batch of one scene of at least 4,194,304 triangles emits and refits
    its top levels with one launch per level over the whole chip: the heap
    words, boxes, codes and ids equal the oracle's, and rays traced through
    the scene (traversal nodes refitted by those launches) equal the
    oracle's trace on its heap.  dup_grid: clustered centroids, so the big
    nodes form long unbalanced chains and leaves at the depth limit."""
    n = (1 << 22) + 12345
    xyz, idx = soup(n, 0.002, 9) if kind == "soup" else dup_grid(n, cells=24, seed=3)
    check_build(ctx, oracle, xyz, idx, 23)
    s = ctx.add_scene(xyz, idx, 23)
    ctx.build(s)
    ref = oracle.build_scene(xyz, idx, 23)
    lo, hi = xyz.min(0), xyz.max(0)
    c = (lo + hi) / 2
    cam = rtcore.camera_look_at(tuple(c + [0.3, 0.4, -1.6]), tuple(c), (0, 1, 0), 50.0, 1.0)
    rays = oracle.raygen_pinhole(cam.as_array(), 96, 96)
    got = ctx.trace(s, rays)
    want = oracle.trace(ref["nodes"], ref["ids"], xyz, idx, rays, nthreads=16)
    assert (want["tri"] >= 0).mean() > 0.2
    assert np.array_equal(got, want)
    ctx.scene_destroy(s)


def test_empty_scene(ctx):
    s = ctx.scene_create(16, 16, 8)
    ctx.scene_upload(s, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    ctx.build(s)
    import rtcore

    rays = np.zeros(4, rtcore.RAY_DTYPE)
    rays["d"] = [0, 0, 1]
    rays["tmax"] = 5.0
    rays["tri"] = -1
    h = ctx.trace(s, rays)
    assert np.all(h["tri"] == -1) and np.all(h["t"] == 5.0)
    ctx.scene_destroy(s)


def test_rebuild_is_deterministic_and_transform_exact(ctx, oracle, rtcore):
    xyz, idx = soup(100000, seed=12)
    s = ctx.add_scene(xyz, idx, 20)
    try:
        ctx.build(s)
        first = ctx.read(s)
        ctx.build(s)
        again = ctx.read(s)
        reach = oracle.reachable(first[0])
        assert np.array_equal(first[0][reach], again[0][reach])
        assert np.array_equal(first[2], again[2])
        for frame in (1, 37):
            m = rtcore.rotation_about((0.3, 1.0, 0.2), float(frame), (0.5, 0.5, 0.5))
            ctx.scene_transform(s, m)
            ctx.build(s)
            nodes, codes, ids, box = ctx.read(s)
            f = np.float32
            d = (xyz - m[9:12]).astype(f)
            R = m[:9].reshape(3, 3)
            rot = np.stack([((((R[r, 0] * d[:, 0]).astype(f) + (R[r, 1] * d[:, 1]).astype(f)).astype(f)
                              + (R[r, 2] * d[:, 2]).astype(f)).astype(f) + m[9 + r]).astype(f)
                            for r in range(3)], 1)
            ref = oracle.build_scene(rot, idx, 20)
            assert np.array_equal(box, ref["scene"])
            assert np.array_equal(codes, ref["codes"]) and np.array_equal(ids, ref["ids"])
            rr = oracle.reachable(ref["nodes"])
            assert np.array_equal(nodes[rr]["prim_count"], ref["nodes"][rr]["prim_count"])
            assert np.all(nodes[rr]["mn"] == ref["nodes"][rr]["mn"])
    finally:
        ctx.scene_destroy(s)


def test_upload_rejects_bad_indices(ctx, rtcore):
    s = ctx.scene_create(4, 4, 6)
    with pytest.raises(rtcore.RtError) as e:
        ctx.scene_upload(s, np.zeros((3, 3), np.float32), np.array([[0, 1, 3]], np.uint32))
    assert e.value.code == rtcore.RT_ERR_ARG
    with pytest.raises(rtcore.RtError) as e:
        ctx.scene_upload(s, np.zeros((9, 3), np.float32), np.zeros((5, 3), np.uint32))
    assert e.value.code == rtcore.RT_ERR_CAPACITY
    ctx.scene_destroy(s)


def test_batch_of_scenes_matches_per_scene_oracle(ctx, oracle, rtcore):
    """Scenes of uneven sizes (and depths) built by one launch sequence."""
    rng = np.random.default_rng(77)
    sizes = [0, 1, 700, 5000, 30000, 123457, 2, 513, 100000]
    depths = [8, 6, 12, 16, 20, 20, 4, 10, 14]
    scenes = []
    for i, n in enumerate(sizes):
        if n == 0:
            scenes.append((np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)))
        elif i % 2:
            scenes.append(soup(n, sigma=0.01, seed=100 + i))
        else:
            scenes.append(rtcore.gen_object_mix(n, seed=200 + i))
    caps = [max(1, n + int(rng.integers(0, 50))) for n in sizes]
    vcaps = [max(1, len(x) + 17) for x, _ in scenes]
    handles = ctx.batch_create(caps, vcaps, depths)
    print(f"batch of {len(handles)} scenes", flush=True)
    for h, (xyz, idx) in zip(handles, scenes):
        ctx.scene_upload(h, xyz, idx)
    ctx.build(handles[3])
    for h, (xyz, idx), d in zip(handles, scenes, depths):
        nodes, codes, ids, box = ctx.read(h)
        ref = oracle.build_scene(xyz, idx, d)
        assert np.array_equal(codes, ref["codes"]) and np.array_equal(ids, ref["ids"])
        if len(idx) == 0:
            continue
        assert np.array_equal(box, ref["scene"])
        reach = oracle.reachable(ref["nodes"])
        assert np.array_equal(nodes[reach]["prim_count"], ref["nodes"][reach]["prim_count"])
        assert np.array_equal(nodes[reach]["prim_pos"], ref["nodes"][reach]["prim_pos"])
        assert np.all(nodes[reach]["mn"] == ref["nodes"][reach]["mn"])
        assert np.all(nodes[reach]["mx"] == ref["nodes"][reach]["mx"])
        # traversal inside the batch uses the scene's own root and local ids
        rays = np.zeros(512, rtcore.RAY_DTYPE)
        c = xyz.mean(0)
        rays["o"] = c + rng.normal(size=(512, 3)).astype(np.float32) * 3
        dvec = (c - rays["o"]) + rng.normal(size=(512, 3)).astype(np.float32) * 0.3
        rays["d"] = (dvec / np.linalg.norm(dvec, axis=1, keepdims=True)).astype(np.float32)
        rays["tmax"] = 1e16
        rays["tri"] = -1
        g = ctx.trace(h, rays)
        r = oracle.trace(ref["nodes"], ref["ids"], xyz, idx, rays)
        assert np.array_equal(g, r)
    # rebuilding after a resize of one scene keeps the others intact
    xyz, idx = scenes[2][0], scenes[2][1][::-1][:650].copy()  # fewer, reordered triangles
    ctx.scene_upload(handles[2], xyz, idx)
    ctx.build(handles[0])
    nodes, codes, ids, box = ctx.read(handles[2])
    ref = oracle.build_scene(xyz, idx, depths[2])
    assert np.array_equal(codes, ref["codes"]) and np.array_equal(ids, ref["ids"])
    nodes, codes, ids, box = ctx.read(handles[5])
    ref = oracle.build_scene(*scenes[5], depths[5])
    assert np.array_equal(ids, ref["ids"])
    for h in handles:
        ctx.scene_destroy(h)


def test_build_loaded_obj_model(ctx, oracle, rtcore, tmp_path):
    """A model read through rt_load_model (the reference's OBJ path and
    normalisation) builds bit for bit like the oracle on the same arrays."""
    xyz, idx = rtcore.gen_noisy_sphere(129, 128, 0.05, 7)
    lines = [f"v {x:.9g} {y:.9g} {z:.9g}" for x, y, z in xyz * 3.0 + [5.0, -2.0, 1.0]]
    lines += [f"f {a + 1} {b + 1} {c + 1}" for a, b, c in idx]
    path = tmp_path / "sphere.obj"
    path.write_text("\n".join(lines) + "\n")
    mx, mi = rtcore.load_model(str(path))
    assert len(mi) == len(idx)
    assert abs(np.sqrt((mx.astype(np.float64) ** 2).sum(1)).max() - 0.49) < 1e-6
    check_build(ctx, oracle, mx, mi, 18, check_partition_mode=True)


def test_batch_above_the_large_tile_threshold(ctx, oracle, rtcore):
    """A batch of 1.2 M triangles sorts with the large tiles, scene pass
    included; every scene's codes, ids and reachable nodes stay bit-exact."""
    sizes = [600011, 499999, 100003, 7]
    depths = [21, 21, 18, 4]
    scenes = [soup(n, sigma=0.005, seed=300 + i) for i, n in enumerate(sizes)]
    handles = ctx.batch_create(sizes, [len(x) for x, _ in scenes], depths)
    for h, (xyz, idx) in zip(handles, scenes):
        ctx.scene_upload(h, xyz, idx)
    ctx.build(handles[0])
    for h, (xyz, idx), d in zip(handles, scenes, depths):
        nodes, codes, ids, box = ctx.read(h)
        ref = oracle.build_scene(xyz, idx, d)
        assert np.array_equal(codes, ref["codes"]) and np.array_equal(ids, ref["ids"]), len(idx)
        reach = oracle.reachable(ref["nodes"])
        assert np.array_equal(nodes[reach]["prim_count"], ref["nodes"][reach]["prim_count"])
        assert np.array_equal(nodes[reach]["prim_pos"], ref["nodes"][reach]["prim_pos"])
    for h in handles:
        ctx.scene_destroy(h)
