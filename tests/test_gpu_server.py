"""Render server on the GPU: every stream's accumulated image equals the same
passes rendered alone with rt_render_jobs (whole image, one job per launch),
whatever mix of requests shared the server's launches and however the tiles
were split over steps; scenes evicted under a small budget and rebuilt give
the same images."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool(rtcore):
    return {
        1: (rtcore.gen_object_mix(5000, seed=1), 16),
        2: (rtcore.gen_triangle_soup(30000, 0.01, seed=2), 18),
        3: (rtcore.gen_torus(), 18),
    }


def view_cam(rtcore, xyz, view, w, h):
    c = xyz.mean(0)
    off = np.array([0.3 + 0.2 * view, 1.0, -3.0 - 0.5 * view], np.float32)
    return rtcore.camera_look_at(tuple(c + off), tuple(c), (0, 1, 0), 60.0, w / h)


def alone(ctx, rtcore, xyz, idx, depth, passes):
    """Renders the passes of one stream, each as its own whole-image launch."""
    s = ctx.add_scene(xyz, idx, depth)
    ctx.build(s)
    w, h = passes[0]["w"], passes[0]["h"]
    acc = ctx.alloc(16 * w * h)
    ctx.zero(acc, 16 * w * h)
    for f, p in enumerate(passes):
        ctx.render_jobs([ctx.job(s, p["cam"], w, h, acc, p["mode"], p["spp"], p["bounces"], 0.7, 1.0, 0.1,
                                 lights=p["lights"], seed=p["seed"], frame=f)], persistent=False,
                        want_stats=False)
    img = ctx.read_accum(acc, w, h)
    ctx.free(acc)
    ctx.scene_destroy(s)
    return img


def test_streams_equal_alone_in_any_mix(ctx, rtcore, pool):
    srv = rtcore.Server(ctx, 8 << 30, max_inflight=6, tiles_per_step=7)
    for k, ((xyz, idx), depth) in pool.items():
        srv.add_scene(k, xyz, idx, depth)
    # (scene, view, w, h, mode, spp, bounces) per request; stream (1, 0) gets three passes
    reqs = [(1, 0, 48, 40, rtcore.MODE_GI, 2, 3), (2, 0, 64, 32, rtcore.MODE_DIRECT, 1, 0),
            (3, 1, 40, 40, rtcore.MODE_GI, 1, 2), (1, 0, 48, 40, rtcore.MODE_GI, 2, 3),
            (2, 2, 33, 17, rtcore.MODE_GI, 3, 4), (1, 0, 48, 40, rtcore.MODE_GI, 2, 3),
            (3, 1, 40, 40, rtcore.MODE_DIRECT, 2, 0)]
    streams = {}
    arrivals = []
    for n, (k, view, w, h, mode, spp, b) in enumerate(reqs):
        xyz = pool[k][0][0]
        cam = view_cam(rtcore, xyz, view, w, h)
        lights = [tuple(xyz.mean(0) + [0, 3, 0])]
        srv.submit(k, view, cam, w, h, mode, spp, b, 0.7, 1.0, 0.1, lights=lights, seed=9, arrival=0.0)
        arrivals.append(0.0)
        streams.setdefault((k, view, w, h, mode), []).append(
            dict(w=w, h=h, cam=cam, mode=mode, spp=spp, bounces=b, lights=lights, seed=9))
    res, _ = srv.run(arrivals)
    assert len(res) == len(reqs) and all(r["status"] == rtcore.RT_OK for r in res)
    st = srv.status()
    assert st["steps"] > len(reqs) and st["builds"] == 3 and st["rays"] > 0
    for key, passes in streams.items():
        got, n = srv.read_stream(*key)
        assert n == sum(p["spp"] for p in passes)
        (xyz, idx), depth = pool[key[0]]
        want = alone(ctx, rtcore, xyz, idx, depth, passes)
        assert np.array_equal(got, want), key
    srv.close()


def test_eviction_and_rebuild_give_same_images(ctx, rtcore, pool):
    probe = rtcore.Server(ctx, 8 << 30, 4, 64)
    (xyz2, idx2), d2 = pool[2]
    probe.add_scene(2, xyz2, idx2, d2)
    probe.submit(2, 0, view_cam(rtcore, xyz2, 0, 32, 32), 32, 32, rtcore.MODE_GI, 1, 2, arrival=0.0)
    probe.run([0.0])
    budget = probe.status()["bytes_used"] + 2 * 16 * 32 * 32  # the largest scene + small streams
    probe.close()
    srv = rtcore.Server(ctx, budget, 4, 64)
    for k, ((xyz, idx), depth) in pool.items():
        srv.add_scene(k, xyz, idx, depth)
    t = 0.0
    for n, k in enumerate([1, 2, 3, 2, 1]):
        (xyz, idx), depth = pool[k]
        cam = view_cam(rtcore, xyz, n, 32, 32)
        srv.submit(k, n, cam, 32, 32, rtcore.MODE_GI, 2, 2, seed=4, arrival=t)
        res, t = srv.run([t], now=t)
        assert len(res) == 1 and res[0]["status"] == rtcore.RT_OK
        got, _ = srv.read_stream(k, n, 32, 32, rtcore.MODE_GI)
        want = alone(ctx, rtcore, xyz, idx, depth,
                     [dict(w=32, h=32, cam=cam, mode=rtcore.MODE_GI, spp=2, bounces=2, lights=None, seed=4)])
        assert np.array_equal(got, want), (n, k)
        assert srv.status()["bytes_used"] <= budget
    st = srv.status()
    assert st["evictions"] > 0 and st["builds"] >= 4
    srv.close()


def test_served_workload_shape_at_scale(ctx, rtcore):
    """The served workload's shape at a tenth of its pool and a fiftieth of
    its requests: 60 object-mix scenes of 1k..500k triangles (log-uniform),
    240 requests arriving at 200 per second (Poisson), 512^2 or 1920x1080,
    direct or 2-bounce GI at 1 spp, four views per scene so that streams
    accumulate progressively, a 200 GB budget, the 1 ms batch window and
    centre-first tiles of the bench's server.  Every request completes, every
    stream holds as many passes as it was sent, latencies are finite, and one
    stream of each kind (direct 512^2, direct 1080p, GI 512^2, GI 1080p) equals
    its passes rendered alone."""
    import math
    rng = np.random.default_rng(56)
    npool, nreq = 60, 240
    sizes = np.exp(rng.uniform(math.log(1000), math.log(500000), npool)).astype(np.int64)
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, nreq))
    scene_of = rng.integers(0, npool, nreq)
    big = rng.random(nreq) < 0.5
    gi = rng.random(nreq) < 0.5
    view_of = rng.integers(0, 4, nreq)
    pool = {int(k): rtcore.gen_object_mix(int(sizes[k]), seed=7000 + int(k)) for k in np.unique(scene_of)}
    depth = {k: int(min(24, max(8, math.ceil(math.log2(max(len(i), 2))) + 2))) for k, (_, i) in pool.items()}
    srv = rtcore.Server(ctx, 200 << 30, max_inflight=32, tiles_per_step=32768, tile_order=rtcore.TILES_CENTRE,
                        batch_window=0.001)
    for k, (x, i) in pool.items():
        srv.add_scene(k, x, i, depth[k])
    streams = {}
    for r in range(nreq):
        k, v = int(scene_of[r]), int(view_of[r])
        xyz = pool[k][0]
        w, h = (1920, 1080) if big[r] else (512, 512)
        lo, hi = xyz.min(0), xyz.max(0)
        c = 0.5 * (lo + hi)
        a = math.radians(30.0 + 90.0 * v)
        pos = c + 1.6 * 0.5 * float(np.linalg.norm(hi - lo)) * np.array([math.cos(a), 0.45, math.sin(a)], np.float32)
        cam = rtcore.camera_look_at(tuple(pos), tuple(c), (0, 1, 0), 60.0, w / h)
        if gi[r]:
            mode, p = rtcore.MODE_GI, dict(spp=1, bounces=2, albedo=0.7, sky=1.0, ambient=0.0, lights=None)
        else:
            mode, p = rtcore.MODE_DIRECT, dict(spp=1, bounces=0, albedo=0.8, sky=0.0, ambient=0.1,
                                               lights=[tuple(hi + 0.5 * (hi - lo))])
        srv.submit(k, v, cam, w, h, mode, p["spp"], p["bounces"], p["albedo"], p["sky"], p["ambient"],
                   lights=p["lights"], seed=3, arrival=float(arrivals[r]))
        streams.setdefault((k, v, w, h, mode), []).append(dict(p, cam=cam))
    res, t_end = srv.run(list(arrivals))
    st = srv.status()
    assert len(res) == nreq and all(q["status"] == rtcore.RT_OK for q in res)
    lat = np.array([q["completed"] - q["arrival"] for q in res])
    print(f"{nreq} requests over {len(pool)} scenes ({int(sizes[np.unique(scene_of)].sum())} tris), "
          f"{len(streams)} streams: p50 {np.percentile(lat, 50) * 1e3:.2f} ms, p99 {np.percentile(lat, 99) * 1e3:.2f} ms, "
          f"{st['steps']} steps, {st['builds']} builds, {st['rays'] / st['device_s'] / 1e6:.0f} Mrays/s of device time",
          flush=True)
    assert np.all(np.isfinite(lat)) and np.all(lat >= 0) and st["builds"] == len(pool)
    checked = set()
    for key, passes in sorted(streams.items(), key=lambda kv: -len(kv[1])):
        k, v, w, h, mode = key
        got, n = srv.read_stream(*key)
        assert n == len(passes), key
        kind = (w, mode)
        if kind in checked:
            continue
        checked.add(kind)
        x, i = pool[k]
        s = ctx.add_scene(x, i, depth[k])
        ctx.build(s)
        acc = ctx.alloc(16 * w * h)
        ctx.zero(acc, 16 * w * h)
        for f, p in enumerate(passes):
            ctx.render_jobs([ctx.job(s, p["cam"], w, h, acc, mode, p["spp"], p["bounces"], p["albedo"], p["sky"],
                                     p["ambient"], lights=p["lights"], seed=3, frame=f)], persistent=False,
                            want_stats=False)
        want = ctx.read_accum(acc, w, h)
        ctx.free(acc)
        ctx.scene_destroy(s)
        assert np.array_equal(got, want), key
    assert len(checked) == 4
    srv.close()
