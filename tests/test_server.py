"""Render-server scheduling on the CPU (rt_server_* in dry-run mode: no device
work, a step takes its estimated rays / dry_rate seconds).  Covers admission
order, the memory budget with least-recently-used eviction, rejection of
requests that can never fit, one request in flight per stream with
successive frame indices, continuous batching (a short request is not held
behind a long one), and argument errors."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def scenes(rtcore):
    small = rtcore.gen_torus(nu=32, nv=16)              # 1,024 triangles
    mid = rtcore.gen_object_mix(20000, seed=3)          # ~20k
    big = rtcore.gen_triangle_soup(120000, 0.01, seed=4)
    return {1: small, 2: mid, 3: big}


def cam(rtcore):
    return rtcore.camera_look_at((0, 0.5, -3), (0, 0, 0), (0, 1, 0), 60.0, 1.0)


def make(rtcore, scenes, budget, inflight=8, tiles=256, rate=1e9):
    s = rtcore.Server(None, budget, inflight, tiles, dry_run=True, dry_rate=rate)
    for k, (xyz, idx) in scenes.items():
        s.add_scene(k, xyz, idx, 18)
    return s


def test_all_requests_complete_in_order_of_arrival(rtcore, scenes):
    s = make(rtcore, scenes, 1 << 34)
    c = cam(rtcore)
    arrivals = [0.001 * k for k in range(30)]
    ids = [s.submit(1 + k % 3, k, c, 64, 64, rtcore.MODE_GI, 2, 3, arrival=t)
           for k, t in enumerate(arrivals)]
    res, t_end = s.run(arrivals)
    assert sorted(r["id"] for r in res) == ids
    for r in res:
        assert r["status"] == rtcore.RT_OK
        assert r["arrival"] <= r["admitted"] < r["completed"] <= t_end
        assert r["rays"] == 64 * 64 * 2 * (1 + 3)  # estimate: spp x (1 + bounces) per pixel
    admitted = [r["admitted"] for r in sorted(res, key=lambda r: r["id"])]
    assert admitted == sorted(admitted)  # admission in arrival order
    st = s.status()
    assert st["completed"] == 30 and st["rejected"] == 0 and st["builds"] == 3
    assert st["evictions"] == 0 and st["pending"] == 0 and st["inflight"] == 0
    assert abs(st["device_s"] - st["rays"] / 1e9) < 1e-9


def test_budget_evicts_least_recently_used(rtcore, scenes):
    # learn the footprints with an unbounded server
    probe = make(rtcore, scenes, 1 << 40)
    c = cam(rtcore)
    sizes = {}
    for k in (1, 2, 3):
        before = probe.status()["bytes_used"]
        probe.submit(k, 0, c, 32, 32, rtcore.MODE_DIRECT, 1, 0, lights=[(0, 3, 0)])
        probe.run([0.0])
        sizes[k] = probe.status()["bytes_used"] - before - 16 * 32 * 32
    assert sizes[1] < sizes[2] < sizes[3]
    # room for scene 3 plus one of the others, not all three
    budget = sizes[3] + sizes[2] + 4 * 16 * 32 * 32
    s = make(rtcore, scenes, budget)
    seq = [1, 2, 3, 1, 3, 2, 2, 1]
    t = 0.0
    for k in seq:
        s.submit(k, 0, c, 32, 32, rtcore.MODE_DIRECT, 1, 0, lights=[(0, 3, 0)], arrival=t)
        res, t = s.run([t], now=t)
        assert len(res) == 1 and res[0]["status"] == rtcore.RT_OK
        st = s.status()
        assert st["bytes_used"] <= budget
    st = s.status()
    assert st["evictions"] > 0
    assert 3 < st["builds"] < len(seq)  # hits on resident scenes skip the build


def footprint(rtcore, scenes, key, w=16, h=16):
    probe = make(rtcore, scenes, 1 << 40)
    probe.submit(key, 0, cam(rtcore), w, h, arrival=0.0)
    probe.run([0.0])
    return probe.status()["bytes_used"]


def test_request_that_never_fits_is_rejected(rtcore, scenes):
    budget = footprint(rtcore, scenes, 1)  # exactly the torus + one 16 x 16 stream
    assert budget < footprint(rtcore, scenes, 3)
    s = make(rtcore, scenes, budget)
    c = cam(rtcore)
    a = s.submit(3, 0, c, 16, 16, arrival=0.0)
    b = s.submit(1, 0, c, 16, 16, arrival=0.0)
    res, _ = s.run([0.0, 0.0])
    by_id = {r["id"]: r for r in res}
    assert by_id[a]["status"] == rtcore.RT_ERR_BUDGET
    assert by_id[b]["status"] == rtcore.RT_OK
    st = s.status()
    assert st["rejected"] == 1 and st["completed"] == 1 and st["bytes_used"] == budget


def test_stream_passes_are_serialised_and_numbered(rtcore, scenes):
    s = make(rtcore, scenes, 1 << 34, inflight=8, tiles=8)
    c = cam(rtcore)
    ids = [s.submit(2, 7, c, 48, 48, rtcore.MODE_GI, 3, 2, arrival=0.0) for _ in range(4)]
    other = s.submit(1, 0, c, 48, 48, rtcore.MODE_GI, 3, 2, arrival=0.0)
    seen_inflight = []
    t = 0.0
    res = []
    for _ in range(10000):
        st = s.status()
        if st["pending"] == 0 and st["inflight"] == 0:
            break
        seen_inflight.append(st["inflight"])
        t += s.step(t)
        res += s.poll()
    by_id = {r["id"]: r for r in res}
    frames = [by_id[i]["frame"] for i in ids]
    assert frames == [0, 1, 2, 3]
    assert [by_id[i]["samples"] for i in ids] == [3, 6, 9, 12]
    done = [by_id[i]["completed"] for i in ids]
    starts = [by_id[i]["admitted"] for i in ids]
    for k in range(3):
        assert starts[k + 1] >= done[k]  # one pass of a stream in flight at a time
    # the other stream was not held behind the busy one
    assert by_id[other]["completed"] <= done[0] + 1e-12
    assert max(seen_inflight) <= 2
    img, n = s.read_stream(2, 7, 48, 48, rtcore.MODE_GI)
    assert n == 12 and img.shape == (48, 48, 4)


def test_short_request_not_held_behind_long_one(rtcore, scenes):
    s = make(rtcore, scenes, 1 << 34, inflight=4, tiles=64)
    c = cam(rtcore)
    long_id = s.submit(3, 0, c, 1024, 1024, rtcore.MODE_GI, 4, 4, arrival=0.0)
    short_id = s.submit(1, 0, c, 64, 64, rtcore.MODE_GI, 1, 1, arrival=0.0001)
    res, _ = s.run([0.0, 0.0001])
    by_id = {r["id"]: r for r in res}
    assert by_id[short_id]["completed"] < by_id[long_id]["completed"]
    # the short one's latency is a small fraction of the long one's
    lat_short = by_id[short_id]["completed"] - by_id[short_id]["arrival"]
    lat_long = by_id[long_id]["completed"] - by_id[long_id]["arrival"]
    assert lat_short < 0.1 * lat_long


def test_inflight_limit_and_tile_budget(rtcore, scenes):
    s = make(rtcore, scenes, 1 << 34, inflight=3, tiles=10)
    c = cam(rtcore)
    for k in range(9):
        s.submit(1 + k % 3, k, c, 64, 64, rtcore.MODE_DIRECT, 1, 0, arrival=0.0)
    peak = 0
    t = 0.0
    for _ in range(10000):
        st = s.status()
        if st["pending"] == 0 and st["inflight"] == 0:
            break
        t += s.step(t)
        peak = max(peak, s.status()["inflight"])
    assert peak <= 3
    # 9 requests x 32 tiles of 16 x 8 pixels at 10 tiles per step
    assert s.status()["steps"] >= 9 * 32 // 10


def test_argument_errors(rtcore, scenes):
    with pytest.raises(rtcore.RtError):
        rtcore.Server(None, 1 << 30, 4, 64, dry_run=False)  # a real server needs a context
    with pytest.raises(rtcore.RtError):
        rtcore.Server(None, 1 << 30, 0, 64, dry_run=True)
    s = make(rtcore, scenes, 1 << 30)
    c = cam(rtcore)
    xyz, idx = scenes[1]
    with pytest.raises(rtcore.RtError):
        s.add_scene(1, xyz, idx, 18)  # key taken
    bad = idx.copy()
    bad[0, 0] = len(xyz)
    with pytest.raises(rtcore.RtError):
        s.add_scene(9, xyz, bad, 18)
    with pytest.raises(rtcore.RtError):
        s.submit(42, 0, c, 16, 16)  # unknown scene
    with pytest.raises(rtcore.RtError):
        s.submit(1, 0, c, 0, 16)
    with pytest.raises(rtcore.RtError):
        s.submit(1, 0, c, 16, 16, mode=7)
    with pytest.raises(rtcore.RtError):
        s.submit(1, 0, c, 16, 16, lights=np.zeros((9, 3)))
    with pytest.raises(rtcore.RtError):
        s.read_stream(1, 0, 16, 16, rtcore.MODE_GI)  # nothing rendered yet


def test_batch_window_groups_arrivals(rtcore, scenes):
    """With a batch window, an arrived request waits (while nothing is in
    flight) until the window since the oldest arrival ends, so requests that
    arrive within it share one launch; a full batch starts at once; without
    a window each request starts when it arrives."""
    c = cam(rtcore)
    arrivals = [0.0, 0.0004, 0.0008, 0.010, 0.0102]
    out = {}
    for window in (0.0, 0.001):
        s = rtcore.Server(None, 1 << 34, 8, 1 << 20, dry_run=True, dry_rate=1e12, batch_window=window)
        for k, (xyz, idx) in scenes.items():
            s.add_scene(k, xyz, idx, 18)
        for k, t in enumerate(arrivals):
            s.submit(1 + k % 3, k, c, 32, 32, rtcore.MODE_DIRECT, 1, 0, lights=[(0, 3, 0)], arrival=t)
        res, _ = s.run(arrivals)
        assert len(res) == 5 and all(r["status"] == rtcore.RT_OK for r in res)
        out[window] = (s.status()["steps"], sorted((r["arrival"], r["admitted"]) for r in res))
        s.close()
    steps0, adm0 = out[0.0]
    steps1, adm1 = out[0.001]
    assert steps0 == 5 and all(a == t for t, a in adm0)  # each starts on arrival, alone
    assert steps1 == 2  # {0, 0.4, 0.8 ms} then {10, 10.2 ms}
    assert [a for _, a in adm1] == pytest.approx([0.001] * 3 + [0.011] * 2)
    # a full batch (max_inflight arrived requests) does not wait
    s = rtcore.Server(None, 1 << 34, 2, 1 << 20, dry_run=True, dry_rate=1e12, batch_window=0.5)
    for k, (xyz, idx) in scenes.items():
        s.add_scene(k, xyz, idx, 18)
    for k in range(2):
        s.submit(1, k, c, 32, 32, rtcore.MODE_DIRECT, 1, 0, lights=[(0, 3, 0)], arrival=0.0)
    res, t_end = s.run([0.0, 0.0])
    assert len(res) == 2 and all(r["admitted"] == 0.0 for r in res) and t_end < 0.5
    s.close()
