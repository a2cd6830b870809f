"""The SAH BVH on the GPU: rt_bvh_build + rt_scene_select make every
traversal kernel (rt_trace, rt_render_direct, rt_render_whitted,
rt_render_jobs) walk the BVH.  Against the oracle's BVH_IntersectP (pinned by
brute force in tests/test_bvh.py) the hits are expected bit-exact: the same
box and triangle arithmetic in the same visit order.  The triangle-test
counter counts each leaf primitive once (the reference's leaf loop tests a
one-primitive leaf twice, BVH.hlsl:165-169), and a miss reports root 1
where the reference's BVH leaves 0."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = np.float32


def random_rays(oracle, xyz, n, seed, self_ids=0):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, oracle.RAY_DTYPE)
    lo, hi = xyz.min(0), xyz.max(0)
    rays["o"] = (lo + (hi - lo) * rng.uniform(-0.2, 1.2, (n, 3))).astype(F)
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F)
    rays["tmax"] = 1e16
    rays["tri"] = -1
    if self_ids:
        rays["tri"][::7] = rng.integers(0, self_ids, len(rays["tri"][::7]))
    return rays


def compare(g, r, tag):
    hit = r["tri"] >= 0
    exact = {k: bool(np.array_equal(g[k], r[k])) for k in ("tri", "u", "v", "t")}
    print(f"{tag}: rays={len(g)} hits={hit.mean():.3f} exact={exact}", flush=True)
    assert all(exact.values()), tag
    assert np.array_equal(g["root"][hit], r["root"][hit])
    assert (g["root"][~hit] == 1).all()
    assert np.array_equal(g["visited"], r["visited"])


@pytest.mark.parametrize("max_prims,split", [(1, 0), (4, 0), (1, 1), (2, 2)])
def test_bvh_trace_matches_oracle(ctx, oracle, rtcore, max_prims, split):
    xyz, idx = rtcore.gen_object_mix(60000, 21)
    s = ctx.add_scene(xyz, idx, 18)
    info = ctx.bvh_build(s, max_prims, split)
    nodes, ids = ctx.bvh_read(s, info["nodes"])
    hn, hids, _ = rtcore.bvh_build_host(xyz, idx, max_prims, split)
    assert np.array_equal(nodes, hn) and np.array_equal(ids, hids)
    rays = random_rays(oracle, xyz, 100000, 3, self_ids=len(idx))
    g = ctx.trace(s, rays)
    r = oracle.trace_bvh(nodes, ids, xyz, idx, rays, ref_leaf_loop=False, nthreads=8)
    compare(g, r, f"bvh max_prims={max_prims} split={split}")
    ga = ctx.trace(s, rays, flags=rtcore.FLAG_ANYHIT)
    ra = oracle.trace_bvh(nodes, ids, xyz, idx, rays, anyhit=True, ref_leaf_loop=False, nthreads=8)
    assert np.array_equal(ga["tri"], ra["tri"]) and np.array_equal(ga["t"], ra["t"])
    ctx.scene_destroy(s)


def test_bvh_and_lbvh_frames_agree(ctx, oracle, rtcore):
    """The same scene through both structures: the closest hits can differ
    only where two triangles tie exactly in t."""
    xyz, idx = rtcore.gen_triangle_soup(262144, 0.004, 2)
    s = ctx.add_scene(xyz, idx, 20)
    ctx.build(s)
    w, h = 640, 360
    cam = rtcore.camera_look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 60.0, w / h)
    lights = [(0.5, 2.5, -1.0)]
    a = ctx.render_direct(s, cam, w, h, lights, flags=rtcore.FLAG_KEEP_HITS | rtcore.FLAG_COUNT)
    ra, ha, oa = ctx.frame_read(w, h, 1, hits=True)
    info = ctx.bvh_build(s)
    b = ctx.render_direct(s, cam, w, h, lights, flags=rtcore.FLAG_KEEP_HITS | rtcore.FLAG_COUNT)
    rb, hb, ob = ctx.frame_read(w, h, 1, hits=True)
    same = (ha["tri"] == hb["tri"]).mean()
    print(f"soup262k {w}x{h}: lbvh box/ray {a['box_tests'] / a['primary_rays']:.1f}, "
          f"bvh box/ray {b['box_tests'] / b['primary_rays']:.1f}, bvh depth {info['max_depth']}, "
          f"ids equal {same:.6f}", flush=True)
    assert same >= 0.9999
    assert np.array_equal(ha["t"], hb["t"])
    assert (ra != rb).any(1).mean() <= 1e-3
    # back to the LBVH: the first frame again, bit for bit
    ctx.scene_select(s, rtcore.STRUCTURE_LBVH)
    ctx.render_direct(s, cam, w, h, lights, flags=rtcore.FLAG_KEEP_HITS)
    rc, hc, _ = ctx.frame_read(w, h, 1, hits=True)
    assert np.array_equal(rc, ra) and np.array_equal(hc, ha)
    ctx.scene_select(s, rtcore.STRUCTURE_BVH)
    ctx.scene_destroy(s)


def test_bvh_whitted_and_jobs(ctx, oracle, rtcore):
    """Whitted frames and render jobs on a BVH scene against the same on the
    LBVH: identical wherever no exact-t tie is involved."""
    xyz, idx = rtcore.gen_torus(0.4, 0.15, 96, 48, 1e-3, 4)
    s = ctx.add_scene(xyz, idx, 16)
    ctx.build(s)
    w, h = 200, 150
    cam = rtcore.camera_look_at((0.2, -1.4, 0.6), (0, 0, 0), (0, 0, 1), 55.0, w / h)
    kw = dict(bounces=3, switches=rtcore.WHITTED_SHADOWS, env=(0.3, 0.4, 0.5))
    ctx.render_whitted(s, cam, w, h, (0.5, -0.5, 1.0), **kw)
    la = ctx.frame_read(w, h, 1)[0]
    acc = ctx.alloc(w * h * 16)
    ctx.zero(acc, w * h * 16)
    ctx.render_jobs([rtcore.Context.job(s, cam, w, h, acc, spp=2, max_bounces=3)], want_stats=False)
    lj = ctx.read_accum(acc, w, h)
    ctx.bvh_build(s)
    ctx.render_whitted(s, cam, w, h, (0.5, -0.5, 1.0), **kw)
    ba = ctx.frame_read(w, h, 1)[0]
    ctx.zero(acc, w * h * 16)
    ctx.render_jobs([rtcore.Context.job(s, cam, w, h, acc, spp=2, max_bounces=3)], want_stats=False)
    bj = ctx.read_accum(acc, w, h)
    assert (la != ba).any(1).mean() <= 2e-3
    assert (lj != bj).any(-1).mean() <= 2e-3
    ctx.free(acc)
    ctx.scene_destroy(s)


def test_bvh_dropped_by_upload(ctx, rtcore):
    xyz, idx = rtcore.gen_torus(0.4, 0.15, 16, 8, 0.0, 1)
    s = ctx.add_scene(xyz, idx, 12)
    rays = np.zeros(4, rtcore.RAY_DTYPE)
    rays["o"] = (0.0, 0.0, -2.0)
    rays["d"] = (0.0, 0.0, 1.0)
    rays["tmax"] = 1e16
    rays["tri"] = -1
    ctx.bvh_build(s)
    ctx.trace(s, rays)  # traversable without an LBVH build
    ctx.scene_upload(s, xyz, idx)
    with pytest.raises(rtcore.RtError):
        ctx.scene_select(s, rtcore.STRUCTURE_BVH)
    with pytest.raises(rtcore.RtError):  # back on the LBVH, which was never built
        ctx.trace(s, rays)
    ctx.scene_destroy(s)
