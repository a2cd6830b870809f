"""GPU render jobs: diffuse path tracing against the FP64 CPU path tracer,
result invariance across launch modes, mixes and launch shapes, graph replay,
and the autotuner.

GI tolerance: the device computes in fp32 (device sin/cos, IEEE div/sqrt);
with the same PCG streams it follows the same paths as the FP64 oracle except
where a path grazes an edge, so the same-stream image is compared by RMSE of
the per-pixel mean radiance (bar 1e-3 for radiance in [0, 1]; measured below
5e-6, while a wrong jitter, seed or bounce rule gives 0.05 or more) and
>= 99.5 % of pixels within 1e-3; the converged comparison (1024 device
samples against 4096 independent FP64 samples) reports the RMSE and bounds it
by the Monte Carlo noise of both estimates (bar 0.01; measured 0.0016)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mix(ctx, oracle, rtcore):
    xyz, idx = rtcore.gen_object_mix(20000, seed=11)
    s = ctx.add_scene(xyz, idx, 18)
    ctx.build(s)
    ref = oracle.build_scene(xyz, idx, 18)
    c = xyz.mean(0)
    cam = rtcore.camera_look_at(tuple(c + [0.0, 2.0, -9.0]), tuple(c), (0, 1, 0), 60.0, 64 / 48)
    yield s, xyz, idx, ref, cam
    ctx.scene_destroy(s)


def gpu_gi(ctx, rtcore, scene, cam, w, h, spp, seed=5, frame=0, persistent=True):
    acc = ctx.alloc(16 * w * h)
    ctx.zero(acc, 16 * w * h)
    st = ctx.render_jobs([ctx.job(scene, cam, w, h, acc, rtcore.MODE_GI, spp, 4, 0.7, 1.0, seed=seed,
                                  frame=frame)], persistent=persistent)
    img = ctx.read_accum(acc, w, h)
    ctx.free(acc)
    return img, st


def test_gi_same_streams_match_fp64(ctx, oracle, rtcore, mix):
    s, xyz, idx, ref, cam = mix
    W, H, spp = 64, 48, 4
    img, st = gpu_gi(ctx, rtcore, s, cam, W, H, spp)
    assert np.all(img[..., 3] == spp)
    g = img[..., 0] / spp
    r, rays = oracle.render_gi(ref["nodes"], ref["ids"], xyz, idx, cam.as_array(), W, H, spp, 4, 0.7,
                               1.0, 5, nthreads=8)
    r = r / spp
    rmse = float(np.sqrt(np.mean((g - r) ** 2)))
    close = float(np.mean(np.abs(g - r) <= 1e-3))
    print(f"GI same streams: rays gpu={st['primary_rays'] + st['bounce_rays']} cpu={rays} "
          f"RMSE={rmse:.5f} within1e-3={close:.4f}", flush=True)
    assert abs((st["primary_rays"] + st["bounce_rays"]) - rays) <= 0.01 * rays
    assert rmse < 1e-3 and close >= 0.995


def test_gi_converged_against_fp64(ctx, oracle, rtcore, mix):
    s, xyz, idx, ref, cam = mix
    W, H = 32, 24
    img, _ = gpu_gi(ctx, rtcore, s, cam, W, H, 1024, seed=77)
    g = img[..., 0] / 1024
    r, _ = oracle.render_gi(ref["nodes"], ref["ids"], xyz, idx, cam.as_array(), W, H, 4096, 4, 0.7, 1.0,
                            77, first_sample=100000, nthreads=16)
    r = r / 4096
    rmse = float(np.sqrt(np.mean((g - r) ** 2)))
    print(f"GI 1024 spp vs FP64 4096 spp (independent samples): RMSE={rmse:.5f}", flush=True)
    assert rmse < 0.01


def test_jobs_mix_and_launch_mode_invariance(ctx, rtcore):
    """A job's image is the same alone, in a mix of scenes from two batches,
    per-job launches or one persistent launch, under every launch shape."""
    a = rtcore.gen_object_mix(5000, seed=1)
    b = rtcore.gen_triangle_soup(30000, 0.01, seed=2)
    c = rtcore.gen_torus()
    hs = ctx.batch_create([len(a[1]), len(b[1])], [len(a[0]), len(b[0])], [16, 18])
    ctx.scene_upload(hs[0], *a)
    ctx.scene_upload(hs[1], *b)
    ctx.build(hs[0])
    sc = ctx.add_scene(*c, 20)
    ctx.build(sc)
    specs = [(hs[0], a, (48, 40), rtcore.MODE_GI, 3), (hs[1], b, (64, 32), rtcore.MODE_DIRECT, 2),
             (sc, c, (40, 40), rtcore.MODE_GI, 2)]
    bufs, jobs = [], []
    for scene, (xyz, _), (w, h), mode, spp in specs:
        ctr = xyz.mean(0)
        cam = rtcore.camera_look_at(tuple(ctr + [0.3, 1.0, -3.0]), tuple(ctr), (0, 1, 0), 60.0, w / h)
        acc = ctx.alloc(16 * w * h)
        bufs.append((acc, w, h))
        jobs.append(ctx.job(scene, cam, w, h, acc, mode, spp, 3, 0.7, 1.0, 0.1,
                            lights=[tuple(ctr + [0, 3, 0])], seed=11, frame=2))

    def run(job_list, persistent):
        for acc, w, h in bufs:
            ctx.zero(acc, 16 * w * h)
        ctx.render_jobs(job_list, persistent=persistent, want_stats=False)
        return [ctx.read_accum(acc, w, h) for acc, w, h in bufs]

    alone = [run([j], False)[k] for k, j in enumerate(jobs)]
    saved = ctx.tuning()
    try:
        for shape in [(64, 16, 0), (64, 20, 1), (128, 12, 2), (128, 20, 1), (128, 26, 0), (128, 22, 2), (256, 12, 1)]:
            ctx.set_tuning(shape[0], shape[1], 2, shape[2])
            for persistent in (False, True):
                together = run(jobs, persistent)
                for k in range(3):
                    assert np.array_equal(together[k], alone[k]), (shape, persistent, k)
    finally:
        ctx.set_tuning(*saved)
    for acc, _, _ in bufs:
        ctx.free(acc)
    for h in hs + [sc]:
        ctx.scene_destroy(h)


def test_jobs_lane_groups_equal_one_lane(ctx, rtcore):
    """Uncounted job launches of LBVH scenes share a wave's live GI paths out
    to lane groups once at most 32 remain (and direct samples' shadow rays);
    counted launches (RT_FLAG_COUNT) keep one lane per ray.  Both give the same
    accumulations bit for bit, and the same ray counts, for GI at 4 and 12
    bounces and for direct jobs, in LBVH scenes and a scene traced through its
    SAH BVH (which has no heap indices and keeps one lane)."""
    a = rtcore.gen_object_mix(20000, seed=3)
    b = rtcore.gen_triangle_soup(40000, 0.01, seed=4)
    sa = ctx.add_scene(*a, 18)
    sb = ctx.add_scene(*b, 18)
    ctx.build(sa)
    ctx.build(sb)
    sbvh = ctx.add_scene(*a, 18)
    ctx.build(sbvh)
    ctx.bvh_build(sbvh, 2)
    ctx.scene_select(sbvh, rtcore.STRUCTURE_BVH)
    specs = [(sa, a, (96, 64), rtcore.MODE_GI, 2, 4), (sb, b, (80, 48), rtcore.MODE_GI, 1, 12),
             (sa, a, (64, 64), rtcore.MODE_DIRECT, 2, 0), (sb, b, (72, 40), rtcore.MODE_DIRECT, 1, 0),
             (sbvh, a, (48, 40), rtcore.MODE_GI, 2, 4)]
    bufs, jobs = [], []
    for scene, (xyz, _), (w, h), mode, spp, nb in specs:
        ctr = xyz.mean(0)
        cam = rtcore.camera_look_at(tuple(ctr + [0.4, 1.5, -4.0]), tuple(ctr), (0, 1, 0), 55.0, w / h)
        acc = ctx.alloc(16 * w * h)
        bufs.append((acc, w, h))
        jobs.append(ctx.job(scene, cam, w, h, acc, mode, spp, nb, 0.7, 1.0, 0.1,
                            lights=[tuple(ctr + [0, 3, 0]), tuple(ctr + [2, 1, -3])], seed=7, frame=1))

    def run(flags, persistent):
        for acc, w, h in bufs:
            ctx.zero(acc, 16 * w * h)
        st = ctx.render_jobs(jobs, persistent=persistent, flags=flags)
        return st, [ctx.read_accum(acc, w, h) for acc, w, h in bufs]

    saved = ctx.tuning()
    try:
        for shape in [(128, 20, 2), (128, 22, 0), (64, 20, 1)]:
            ctx.set_tuning(shape[0], shape[1], 4, shape[2])
            st1, one = run(rtcore.FLAG_COUNT, False)
            assert st1["box_tests"] > 0
            for persistent in (False, True):
                st2, grp = run(0, persistent)
                assert st2["box_tests"] == 0  # counts come with RT_FLAG_COUNT only
                for k in ("primary_rays", "shadow_rays", "bounce_rays"):
                    assert st1[k] == st2[k], (shape, persistent, k)
                for k in range(len(bufs)):
                    assert np.array_equal(one[k], grp[k]), (shape, persistent, k)
    finally:
        ctx.set_tuning(*saved)
    for acc, _, _ in bufs:
        ctx.free(acc)
    for h in (sa, sb, sbvh):
        ctx.scene_destroy(h)


def test_trace_parity_under_every_launch_shape(ctx, oracle, rtcore):
    xyz, idx = rtcore.gen_triangle_soup(100000, 0.006, seed=3)
    s = ctx.add_scene(xyz, idx, 20)
    ctx.build(s)
    ref = oracle.build_scene(xyz, idx, 20)
    cam = rtcore.camera_look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 60.0, 1.0)
    rays = oracle.raygen_pinhole(cam.as_array(), 160, 160)
    r = oracle.trace(ref["nodes"], ref["ids"], xyz, idx, rays, nthreads=8)
    saved = ctx.tuning()
    try:
        for form in (0, 1, 2):
            for block, lds in [(64, 16), (64, 20), (128, 12), (128, 16), (128, 20), (128, 22), (128, 26), (256, 12)]:
                ctx.set_tuning(block, lds, 4, form)
                assert np.array_equal(ctx.trace(s, rays), r), (block, lds, form)
                st = ctx.render_direct(s, cam, 160, 160, [(0.5, 2.5, -1.0)], flags=rtcore.FLAG_KEEP_HITS)
                _, hits, _ = ctx.frame_read(160, 160, 1, hits=True)
                assert np.array_equal(hits, r), (block, lds, form)
    finally:
        ctx.set_tuning(*saved)
    ctx.scene_destroy(s)


def test_graph_replay_equals_eager(ctx, rtcore):
    xyz, idx = rtcore.gen_object_mix(8000, seed=21)
    hs = ctx.batch_create([len(idx), len(idx)], [len(xyz), len(xyz)], [16, 16])
    for h in hs:
        ctx.scene_upload(h, xyz, idx)
    ctx.build(hs[0])
    c = xyz.mean(0)
    cam = rtcore.camera_look_at(tuple(c + [0, 1, -6]), tuple(c), (0, 1, 0), 60.0, 1.0)
    W = H = 32
    accs = [ctx.alloc(16 * W * H) for _ in hs]
    jobs = [ctx.job(h, cam, W, H, a, rtcore.MODE_GI, 2, 4, 0.7, 1.0, seed=3, frame=10)
            for h, a in zip(hs, accs)]
    for a in accs:
        ctx.zero(a, 16 * W * H)
    g = ctx.graph_capture(hs[0], jobs, rebuild=True)
    for _ in range(3):  # replays with eager work and syncs in between
        ctx.graph_launch(g)
        ctx.sync()
        ctx.build(hs[1])
        ctx.sync()
    replayed = [ctx.read_accum(a, W, H) for a in accs]
    for a in accs:
        ctx.zero(a, 16 * W * H)
    for f in range(3):
        eager = [ctx.job(h, cam, W, H, a, rtcore.MODE_GI, 2, 4, 0.7, 1.0, seed=3, frame=10 + f)
                 for h, a in zip(hs, accs)]
        ctx.render_jobs(eager, want_stats=False)
    expect = [ctx.read_accum(a, W, H) for a in accs]
    for r, e in zip(replayed, expect):
        assert np.array_equal(r, e)
        assert np.all(r[..., 3] == 6)
    ctx.graph_destroy(g)
    for a in accs:
        ctx.free(a)
    for h in hs:
        ctx.scene_destroy(h)


def test_graph_refuses_replay_after_its_batch_changed(ctx, rtcore):
    """A captured graph holds its batch: destroying the scene handle, building
    or switching to a BVH, or a new upload dropping the BVH makes
    rt_graph_launch fail with RT_ERR_STATE before anything runs; a fresh
    capture works again; destroying the graph releases the batch."""
    xyz, idx = rtcore.gen_object_mix(3000, seed=31)
    c = xyz.mean(0)
    cam = rtcore.camera_look_at(tuple(c + [0, 1, -6]), tuple(c), (0, 1, 0), 60.0, 1.0)
    W = H = 16
    acc = ctx.alloc(16 * W * H)
    ctx.zero(acc, 16 * W * H)
    base = ctx.memory_in_use()
    s = ctx.add_scene(xyz, idx, 14)
    ctx.build(s)
    scene_bytes = ctx.memory_in_use() - base

    def capture():
        return ctx.graph_capture(s, [ctx.job(s, cam, W, H, acc, rtcore.MODE_GI, 1, 2, seed=3)], rebuild=True)

    def refused(g):
        with pytest.raises(rtcore.RtError) as e:
            ctx.graph_launch(g)
        assert e.value.code == rtcore.RT_ERR_STATE

    g = capture()
    ctx.graph_launch(g)
    ctx.sync()
    ctx.bvh_build(s)  # new BVH arrays, and the scene now traverses them
    refused(g)
    ctx.graph_destroy(g)
    g = capture()
    ctx.graph_launch(g)
    ctx.sync()
    ctx.scene_select(s, rtcore.STRUCTURE_LBVH)
    refused(g)
    ctx.graph_destroy(g)
    ctx.bvh_build(s)
    g = capture()
    ctx.scene_upload(s, xyz, idx)  # new vertices drop the BVH the graph reads
    refused(g)
    ctx.graph_destroy(g)
    ctx.build(s)
    g = capture()
    ctx.scene_upload(s, xyz, idx)  # an LBVH graph rebuilds from the new vertices: still valid
    ctx.graph_launch(g)
    ctx.sync()
    ctx.scene_destroy(s)
    refused(g)
    assert ctx.memory_in_use() - base >= scene_bytes  # the graph still holds the batch
    ctx.graph_destroy(g)
    assert ctx.memory_in_use() - base < scene_bytes
    ctx.free(acc)


def test_autotuner_picks_and_caches(ctx, rtcore, tmp_path):
    xyz, idx = rtcore.gen_object_mix(20000, seed=4)
    s = ctx.add_scene(xyz, idx, 18)
    ctx.build(s)
    c = xyz.mean(0)
    cam = rtcore.camera_look_at(tuple(c + [0, 1, -8]), tuple(c), (0, 1, 0), 60.0, 1.0)
    acc = ctx.alloc(16 * 128 * 128)
    ctx.zero(acc, 16 * 128 * 128)
    jobs = [ctx.job(s, cam, 128, 128, acc, rtcore.MODE_GI, 1, 2, 0.7, 1.0)]
    cache = str(tmp_path / "tune.txt")
    saved = ctx.tuning()
    ctx.set_tuning(128, 26, 4, 0, 64)
    best = ctx.autotune(jobs, cache)
    assert best[0] in (64, 128, 256) and 1 <= best[1] <= 26 and best[2] in (2, 4, 8) and best[3] in (0, 1, 2)
    assert ctx.tuning() == best
    assert best[4] == 64  # quads are a direct-frame choice: the search keeps the one set
    assert open(cache).read().split()[-6:] == [str(v) for v in best] + ["0"]
    ctx.set_tuning(128, 26, 4, 0, 64)
    assert ctx.autotune(jobs, cache) == best  # cache hit
    ctx.set_tuning(*saved)
    # the search rendered into scratch buffers, not into the caller's
    assert np.all(ctx.read_accum(acc, 128, 128) == 0)
    ctx.free(acc)
    ctx.scene_destroy(s)


def test_tile_ranges_compose_to_whole_image(ctx, rtcore, mix):
    """Jobs over disjoint tile ranges of one image (one of them empty, one
    running past the end) add up to the whole-image job, in one launch or many."""
    s, xyz, idx, ref, cam = mix
    W, H = 50, 30
    block = ctx.tuning()[0]
    ntiles = ((W + 15) // 16) * ((H + block // 16 - 1) // (block // 16))
    acc = ctx.alloc(16 * W * H)
    ctx.zero(acc, 16 * W * H)
    ctx.render_jobs([ctx.job(s, cam, W, H, acc, rtcore.MODE_GI, 2, 3, seed=8)], want_stats=False)
    whole = ctx.read_accum(acc, W, H)
    cuts = [(0, 3), (3, 0), (3, ntiles // 2), (ntiles // 2, ntiles + 7)]
    for persistent in (True, False):
        ctx.zero(acc, 16 * W * H)
        jobs = [ctx.job(s, cam, W, H, acc, rtcore.MODE_GI, 2, 3, seed=8, tile_begin=b, tile_count=e - b)
                for b, e in cuts if e > b]
        jobs.insert(1, ctx.job(s, cam, W, H, acc, rtcore.MODE_GI, 2, 3, seed=8, tile_begin=ntiles + 5,
                               tile_count=4))  # past the end: renders nothing
        ctx.render_jobs(jobs, persistent=persistent, want_stats=False)
        assert np.array_equal(ctx.read_accum(acc, W, H), whole), persistent
    ctx.free(acc)


def test_centre_tile_order_keeps_the_image(ctx, rtcore, mix):
    """RT_TILES_CENTRE hands a job's tiles out nearest the image centre first:
    the same image as row order, whole or split into tile ranges, in one
    persistent launch or one launch per job, and through the no-allocation
    entry point (which reuses the order made here)."""
    s, xyz, idx, ref, _ = mix
    W, H = 70, 45
    block = ctx.tuning()[0]
    ntiles = ((W + 15) // 16) * ((H + block // 16 - 1) // (block // 16))
    acc = ctx.alloc(16 * W * H)

    def render(jobs, persistent, fn=None):
        ctx.zero(acc, 16 * W * H)
        (fn or ctx.render_jobs)(jobs, persistent=persistent, want_stats=False)
        return ctx.read_accum(acc, W, H)

    def job(order, b=0, c=0, mode=rtcore.MODE_GI):
        return ctx.job(s, ref_cam(), W, H, acc, mode, 2, 3, lights=[tuple(xyz.mean(0) + [0, 4, 0])], seed=12,
                       tile_begin=b, tile_count=c, tile_order=order)

    def ref_cam():
        c = xyz.mean(0)
        return rtcore.camera_look_at(tuple(c + [0.5, 2.0, -7.0]), tuple(c), (0, 1, 0), 60.0, W / H)

    for mode in (rtcore.MODE_GI, rtcore.MODE_DIRECT):
        rows = render([job(rtcore.TILES_ROWS, mode=mode)], True)
        for persistent in (True, False):
            assert np.array_equal(render([job(rtcore.TILES_CENTRE, mode=mode)], persistent), rows)
            cuts = [(0, 5), (5, ntiles // 2), (ntiles // 2, ntiles)]
            parts = [job(rtcore.TILES_CENTRE, b, e - b, mode) for b, e in cuts]
            assert np.array_equal(render(parts, persistent), rows), persistent
            assert np.array_equal(render(parts, persistent, ctx.render_jobs_noalloc), rows)
    ctx.free(acc)


def test_autotune_direct_picks_and_caches(ctx, rtcore, mix, tmp_path):
    s, xyz, idx, ref, cam = mix
    cache = str(tmp_path / "tune_direct.txt")
    saved = ctx.tuning()
    try:
        best = ctx.autotune_direct(s, cam, 96, 64, [(0.0, 5.0, 0.0)], cache_path=cache)
        assert best[0] in (64, 128, 256) and 1 <= best[1] <= 26 and best[2] == saved[2] and best[3] in (0, 1, 2)
        assert best[4] in (0, 64, 256, 1024)  # quads on the heaviest tiles, or none
        assert ctx.tuning() == best
        assert open(cache).read().split()[-6:] == [str(v) for v in best] + ["0"]
        ctx.set_tuning(64, 8, saved[2])
        assert ctx.autotune_direct(s, cam, 96, 64, [(0.0, 5.0, 0.0)], cache_path=cache) == best  # cache hit
        # one cache file shared by the three searches: every record keeps its own hit
        acc = ctx.alloc(16 * 64 * 48)
        jobs = [ctx.job(s, cam, 64, 48, acc, rtcore.MODE_GI, 1, 2, 0.7, 1.0)]
        bj = ctx.autotune(jobs, cache)
        bw = ctx.autotune_whitted(s, cam, 96, 64, (0.0, 5.0, 0.0), bounces=1, cache_path=cache)
        with open(cache, "a") as f:
            f.write("garbage line with words\n" + "direct:short 1 2\n")
        ctx.set_tuning(64, 8, saved[2])
        assert ctx.autotune_direct(s, cam, 96, 64, [(0.0, 5.0, 0.0)], cache_path=cache) == best
        assert ctx.autotune(jobs, cache)[:4] == bj[:4]
        tw, sw = ctx.autotune_whitted(s, cam, 96, 64, (0.0, 5.0, 0.0), bounces=1, cache_path=cache)
        assert tw[:4] == bw[0][:4] and sw == bw[1]
        assert len(open(cache).read().splitlines()) == 5  # hits append nothing
        ctx.free(acc)
    finally:
        ctx.set_tuning(*saved)


def test_batched_object_mix_full_size(ctx, oracle, rtcore):
    """The batched GI workload at its full size: 64 object-mix scenes of
    1,024..200,000 triangles (log-uniform, about 2.4M in all), each at 256^2,
    512^2 or 1024^2, 4 spp x 4 bounces.  Two frames of the whole batch in
    one persistent launch each equal, bit for bit, the same frames rendered
    scene by scene; and a 256^2 scene's frame matches the FP64 oracle
    following the same PCG streams (the same-stream bars above)."""
    import math
    n = 64
    rng = np.random.default_rng(44)
    sizes = np.exp(rng.uniform(math.log(1024), math.log(200000), n)).astype(np.int64)
    res = rng.choice([256, 512, 1024], n)
    scenes = [rtcore.gen_object_mix(int(sizes[k]), seed=1000 + k) for k in range(n)]
    depths = [int(min(24, max(8, math.ceil(math.log2(max(len(i), 2))) + 2))) for _, i in scenes]
    cams = []
    for (x, _), r in zip(scenes, res):
        lo, hi = x.min(0), x.max(0)
        c = 0.5 * (lo + hi)
        rad = 0.5 * float(np.linalg.norm(hi - lo))
        a = math.radians(30.0)
        pos = c + 1.6 * rad * np.array([math.cos(a), 0.45, math.sin(a)], np.float32)
        cams.append(rtcore.camera_look_at(tuple(pos), tuple(c), (0, 1, 0), 60.0, 1.0))
    hs = ctx.batch_create([len(i) for _, i in scenes], [len(x) for x, _ in scenes], depths)
    for h_, (x, i) in zip(hs, scenes):
        ctx.scene_upload(h_, x, i)
    ps = [ctx.add_scene(x, i, d) for (x, i), d in zip(scenes, depths)]
    acc_b = [ctx.alloc(16 * int(r) * int(r)) for r in res]
    acc_p = [ctx.alloc(16 * int(r) * int(r)) for r in res]
    for a, r in zip(acc_b + acc_p, list(res) + list(res)):
        ctx.zero(a, 16 * int(r) * int(r))

    def jobs(handles, accs, frame):
        return [ctx.job(handles[k], cams[k], int(res[k]), int(res[k]), accs[k], rtcore.MODE_GI, 4, 4, 0.7, 1.0,
                        seed=7, frame=frame) for k in range(n)]

    rays_b = rays_p = 0
    for f in range(2):
        ctx.build(hs[0])
        st = ctx.render_jobs(jobs(hs, acc_b, f), persistent=True)
        rays_b += st["primary_rays"] + st["bounce_rays"]
        js = jobs(ps, acc_p, f)
        for k in range(n):
            ctx.build(ps[k])
            st = ctx.render_jobs([js[k]], persistent=False)
            rays_p += st["primary_rays"] + st["bounce_rays"]
    print(f"64 scenes, {int(sizes.sum())} tris, {int((res.astype(np.int64) ** 2).sum())} px: "
          f"{rays_b} rays batched, {rays_p} scene by scene", flush=True)
    assert rays_b == rays_p
    for k in range(n):
        r = int(res[k])
        assert np.array_equal(ctx.read_accum(acc_b[k], r, r), ctx.read_accum(acc_p[k], r, r)), k
    # the smallest 256^2 scene against the FP64 oracle, both frames' streams
    k = min((k for k in range(n) if res[k] == 256), key=lambda k: sizes[k])
    xyz, idx = scenes[k]
    ref = oracle.build_scene(xyz, idx, depths[k])
    want = np.zeros((256, 256))
    for f in range(2):
        want += oracle.render_gi(ref["nodes"], ref["ids"], xyz, idx, cams[k].as_array(), 256, 256, 4, 4, 0.7, 1.0,
                                 7, frame=f, nthreads=16)[0]
    got = ctx.read_accum(acc_b[k], 256, 256)
    assert np.all(got[..., 3] == 8)
    g, w = got[..., 0] / 8, want / 8
    rmse = float(np.sqrt(np.mean((g - w) ** 2)))
    close = float(np.mean(np.abs(g - w) <= 1e-3))
    print(f"scene {k} ({len(idx)} tris, 256^2, 8 spp over two frames): GI vs FP64 same streams "
          f"RMSE {rmse:.5f}, within 1e-3 {close:.4f}", flush=True)
    assert rmse < 1e-3 and close >= 0.995
    for a in acc_b + acc_p:
        ctx.free(a)
    for p in ps:
        ctx.scene_destroy(p)
    for h_ in hs:
        ctx.scene_destroy(h_)
