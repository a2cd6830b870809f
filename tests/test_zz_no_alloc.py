"""rt_render_jobs_noalloc (the render-jobs entry point that allocates no
device memory: job table, tile counter and ray counters in a static device
array) against rt_render_jobs, through the C ABI: two calls in a row, the
second smaller, in both launch modes, must give the same accumulators and
ray counts bit for bit; too many jobs are refused before anything runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(ctx, rtcore, scenes, sizes, accs, modes, frame):
    out = []
    for k, ((s, xyz), (w, h), acc, mode) in enumerate(zip(scenes, sizes, accs, modes)):
        c = xyz.mean(0)
        cam = rtcore.camera_look_at(tuple(c + [0.4, 1.0, -3.5]), tuple(c), (0, 1, 0), 60.0, w / h)
        out.append(ctx.job(s, cam, w, h, acc, mode, 2, 2, 0.7, 1.0, 0.1, lights=[tuple(c + [0, 3, 0])],
                           seed=5 + k, frame=frame))
    return out


def test_noalloc_equals_render_jobs(ctx, rtcore):
    a = rtcore.gen_object_mix(6000, seed=61)
    b = rtcore.gen_triangle_soup(40000, 0.01, seed=62)
    scenes = []
    for xyz, idx in (a, b):
        s = ctx.add_scene(xyz, idx, 18)
        ctx.build(s)
        scenes.append((s, xyz))
    calls = [  # (scenes, sizes, modes): the second call has fewer and smaller jobs
        ([scenes[0], scenes[1], scenes[0]], [(96, 64), (80, 80), (33, 17)],
         [rtcore.MODE_GI, rtcore.MODE_DIRECT, rtcore.MODE_DIRECT]),
        ([scenes[1]], [(40, 24)], [rtcore.MODE_GI]),
    ]
    for persistent in (True, False):
        for n, (sc, sizes, modes) in enumerate(calls):
            accs = {}
            for way in ("plain", "noalloc"):
                bufs = [ctx.alloc(16 * w * h) for w, h in sizes]
                for buf, (w, h) in zip(bufs, sizes):
                    ctx.zero(buf, 16 * w * h)
                jobs = _jobs(ctx, rtcore, sc, sizes, bufs, modes, frame=n)
                fn = ctx.render_jobs if way == "plain" else ctx.render_jobs_noalloc
                st = fn(jobs, persistent=persistent)
                accs[way] = ([ctx.read_accum(buf, w, h) for buf, (w, h) in zip(bufs, sizes)], st)
                for buf in bufs:
                    ctx.free(buf)
            (ga, gs), (na, ns) = accs["plain"], accs["noalloc"]
            for x, y in zip(ga, na):
                assert np.array_equal(x, y), (persistent, n)
            for k in ("primary_rays", "shadow_rays", "bounce_rays"):
                assert gs[k] == ns[k] and gs[k] >= 0, (persistent, n, k)
            assert ns["primary_rays"] == sum(2 * w * h for w, h in sizes)
    for s, _ in scenes:
        ctx.scene_destroy(s)


def test_noalloc_refuses_too_many_jobs(ctx, rtcore):
    xyz, idx = rtcore.gen_torus(nu=16, nv=8)
    s = ctx.add_scene(xyz, idx, 10)
    ctx.build(s)
    acc = ctx.alloc(16 * 4 * 4)
    ctx.zero(acc, 16 * 4 * 4)
    cam = rtcore.camera_look_at((0, 0.5, 3), (0, 0, 0), (0, 1, 0), 60.0, 1.0)
    jobs = [ctx.job(s, cam, 4, 4, acc, rtcore.MODE_DIRECT, 1, 0)] * (rtcore.NOALLOC_MAX_JOBS + 1)
    with pytest.raises(rtcore.RtError) as e:
        ctx.render_jobs_noalloc(jobs)
    assert e.value.code == rtcore.RT_ERR_CAPACITY
    assert np.all(ctx.read_accum(acc, 4, 4) == 0)  # nothing ran
    ctx.free(acc)
    ctx.scene_destroy(s)
