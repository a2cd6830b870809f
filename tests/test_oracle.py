"""Pins of the CPU oracle (oracle/lbvh_oracle.c) -- CPU only.

The reference ships no tests, fixtures or golden vectors for this path, and its
CPU builder cannot be compiled here without stand-in headers, so each oracle
function is pinned by checks that do not restate its own code:
  orc_morton_codes   hand vectors (corners, q = 1024, single-axis bits), the
                     closed-form Part1By2 of MortonCode.hlsl, numpy float32
                     quantisation, code < 2^30
  orc_prim_bounds    numpy float32 min/max/(A+B)+C
  orc_sort_pairs     numpy lexsort on (code, id)
  orc_lbvh_build     hand-built tiny trees (axis bits, leaf rules, depth
                     limit), an independent top-down builder written from the
                     split definition, range/partition invariants, and the
                     in-place-partition mode against the presorted mode
  orc_trace          analytic single-triangle hits, visit counter on known
                     trees, brute force over all triangles (numpy float32)
  orc_raygen_pinhole numpy float32 evaluation of the pixel formula
  orc_render_direct  analytic floor + occluder + light scene
"""
import numpy as np
import pytest

from oracle import HIT_DTYPE, INVALID, NODE_DTYPE, RAY_DTYPE


# --------------------------------------------------------------------- Morton
def _interleave(qx, qy, qz):
    code = 0
    for k in range(10):
        code |= ((qz >> k) & 1) << (3 * k)
        code |= ((qy >> k) & 1) << (3 * k + 1)
        code |= ((qx >> k) & 1) << (3 * k + 2)
    return code


def test_morton_hand_vectors(oracle):
    scene = np.array([0, 0, 0, 1, 1, 1], np.float32)  # 3*min = 0, 3*max = 3
    f = np.float32(1024.0) / np.float32(3.0)

    def cent_for(q):  # centre of quantisation cell q
        return np.float32((q + 0.5) / f)

    cases = {
        (0, 0, 0): 0,
        (1, 0, 0): 1 << 2,
        (0, 1, 0): 1 << 1,
        (0, 0, 1): 1 << 0,
        (2, 0, 0): 1 << 5,
        (0, 0, 512): 1 << 27,
        (512, 0, 0): 1 << 29,
        (1023, 1023, 1023): (1 << 30) - 1,
    }
    cent = np.array([[cent_for(a), cent_for(b), cent_for(c)] for a, b, c in cases], np.float32)
    codes = oracle.morton_codes(cent, scene)
    assert [int(c) for c in codes] == list(cases.values())
    # the maximum corner quantises to q = 1024, whose bit 10 is dropped
    top = oracle.morton_codes(np.array([[3, 3, 3], [3, 0, 0], [0, 3, 1.5]], np.float32), scene)
    assert int(top[0]) == 0 and int(top[1]) == 0
    assert int(top[2]) == _interleave(0, 0, 512)


def test_morton_closed_form_and_quantisation(oracle):
    rng = np.random.default_rng(11)
    q = rng.integers(0, 1025, size=(4000, 3))
    for qx, qy, qz in q[:500]:
        assert oracle.morton_closed_form(int(qx), int(qy), int(qz)) == _interleave(qx & 1023, qy & 1023, qz & 1023)
    scene = np.array([-0.7, 0.2, -3.0, 1.3, 0.9, 4.0], np.float32)
    cent = (rng.random((4000, 3)) * 3 * (scene[3:] - scene[:3]) + 3 * scene[:3]).astype(np.float32)
    codes = oracle.morton_codes(cent, scene)
    m = (np.float32(3.0) * scene[:3]).astype(np.float32)
    M = (np.float32(3.0) * scene[3:]).astype(np.float32)
    f = (np.float32(1024.0) / (M - m)).astype(np.float32)
    s = ((cent - m).astype(np.float32) * f).astype(np.float32)
    qq = s.astype(np.int64)
    exp = [_interleave(a & 1023, b & 1023, c & 1023) for a, b, c in qq]
    assert np.array_equal(codes.astype(np.int64), np.array(exp, np.int64))
    assert int(codes.max()) < (1 << 30)


# --------------------------------------------------------------------- bounds
def test_prim_bounds_numpy(oracle, rtcore):
    xyz, idx = rtcore.gen_torus(seed=5)
    bmin, bmax, cent, scene = oracle.prim_bounds(xyz, idx)
    A, B, Cc = xyz[idx[:, 0]], xyz[idx[:, 1]], xyz[idx[:, 2]]
    assert np.array_equal(bmin, np.minimum(A, np.minimum(B, Cc)))
    assert np.array_equal(bmax, np.maximum(A, np.maximum(B, Cc)))
    assert np.array_equal(cent, ((A + B).astype(np.float32) + Cc).astype(np.float32))
    assert np.array_equal(scene[:3], bmin.min(0)) and np.array_equal(scene[3:], bmax.max(0))


def test_sort_pairs_lexsort(oracle):
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 64, size=5000).astype(np.uint32)  # many duplicates
    ids = rng.permutation(5000).astype(np.int32)
    sc, si = oracle.sort_pairs(codes, ids)
    order = np.lexsort((ids, codes))
    assert np.array_equal(sc, codes[order]) and np.array_equal(si, ids[order])


# ---------------------------------------------------------------------- build
def _boxes(n):
    bmin = np.zeros((n, 3), np.float32) + np.arange(n, dtype=np.float32)[:, None]
    return bmin, bmin + 1


def test_build_tiny_trees(oracle):
    # one primitive: root leaf, count 1, axis 0
    bmin, bmax = _boxes(1)
    nodes, _, _, reached = oracle.lbvh_build(np.array([5], np.uint32), np.array([0], np.int32), 4, bmin, bmax)
    assert reached == 1 and nodes[1]["prim_count"] == (1 << 3) | 4 | 0 and nodes[1]["prim_pos"] == 0
    # two codes differing at bit 29: root split, axis (30-29)%3 = 1; leaves
    # entered under bit 29 get axis (31-29)%3 = 2
    bmin, bmax = _boxes(2)
    nodes, _, _, _ = oracle.lbvh_build(np.array([0, 1 << 29], np.uint32), np.arange(2, dtype=np.int32), 4, bmin, bmax)
    assert nodes[1]["prim_count"] == (2 << 3) | 1
    assert nodes[2]["prim_count"] == (1 << 3) | 4 | 2 and nodes[3]["prim_count"] == (1 << 3) | 4 | 2
    assert nodes[3]["prim_pos"] == 1
    assert np.array_equal(nodes[1]["mn"], [0, 0, 0]) and np.array_equal(nodes[1]["mx"], [2, 2, 2])
    # split at bit 0: internal axis 30 % 3 = 0, leaves (31-0)%3 = 1
    nodes, _, _, _ = oracle.lbvh_build(np.array([6, 7], np.uint32), np.arange(2, dtype=np.int32), 4, bmin, bmax)
    assert nodes[1]["prim_count"] == (2 << 3) | 0
    assert nodes[2]["prim_count"] == (1 << 3) | 4 | 1
    # equal codes: one leaf reached after all 31 bits, axis 31 % 3 = 1
    nodes, _, _, _ = oracle.lbvh_build(np.array([9, 9], np.uint32), np.arange(2, dtype=np.int32), 4, bmin, bmax)
    assert nodes[1]["prim_count"] == (2 << 3) | 4 | 1
    assert nodes[2]["prim_count"] == INVALID
    # depth 2: children of the root are leaves at entry regardless of codes
    bmin, bmax = _boxes(4)
    codes = np.array([0, 1, 1 << 20, (1 << 20) + 3], np.uint32)
    nodes, _, _, reached = oracle.lbvh_build(codes, np.arange(4, dtype=np.int32), 2, bmin, bmax)
    assert reached == 3
    assert nodes[1]["prim_count"] == (4 << 3) | ((30 - 20) % 3)
    assert nodes[2]["prim_count"] == (2 << 3) | 4 | ((31 - 20) % 3)
    assert nodes[3]["prim_count"] == (2 << 3) | 4 | ((31 - 20) % 3) and nodes[3]["prim_pos"] == 2


def independent_build(codes, depth, bmin, bmax, ids):
    """Top-down build from the split definition: a range splits at the first
    index whose highest differing bit (codes[first] ^ codes[last]) is set;
    axis bits from the closed forms (30 - p) % 3, (31 - p_parent) % 3 and 1.
    Written independently of build_recursive's incremental axis."""
    nodes = {}
    leaf_level = 1 << (depth - 1)
    stack = [(1, 0, len(codes), 31)]
    while stack:
        h, first, cnt, pp = stack.pop()
        last = first + cnt - 1
        sel = ids[first:first + cnt]
        box = (bmin[sel].min(0), bmax[sel].max(0))
        if cnt <= 1 or h >= leaf_level or pp == 0:
            nodes[h] = ((cnt << 3) | 4 | ((31 - pp) % 3), first, box)
            continue
        x = int(codes[first]) ^ int(codes[last])
        if x == 0:
            nodes[h] = ((cnt << 3) | 4 | 1, first, box)
            continue
        p = x.bit_length() - 1
        bits = (codes[first:last + 1] >> p) & 1
        s = first + int(np.argmax(bits == 1))
        nodes[h] = ((cnt << 3) | ((30 - p) % 3), first, box)
        stack.append((2 * h, first, s - first, p))
        stack.append((2 * h + 1, s, last - s + 1, p))
    return nodes


@pytest.mark.parametrize("depth,scene", [(20, "torus"), (8, "torus"), (12, "dups")])
def test_build_against_independent_builder(oracle, rtcore, depth, scene):
    if scene == "torus":
        xyz, idx = rtcore.gen_torus(seed=9)
    else:  # many duplicate Morton codes: triangles stacked on a coarse grid
        rng = np.random.default_rng(1)
        c = np.floor(rng.random((3000, 3)) * 6) / 6
        xyz = np.concatenate([c, c + [0.01, 0, 0], c + [0, 0.01, 0]]).astype(np.float32)
        idx = np.stack([np.arange(3000), np.arange(3000) + 3000, np.arange(3000) + 6000], 1).astype(np.uint32)
    b = oracle.build_scene(xyz, idx, depth, presorted=True)
    ref = independent_build(b["codes"], depth, b["bmin"], b["bmax"], b["ids"])
    reach = oracle.reachable(b["nodes"])
    assert sorted(ref) == list(reach)
    for h in reach:
        pc, pos, (lo, hi) = ref[int(h)]
        n = b["nodes"][h]
        assert int(n["prim_count"]) == pc and int(n["prim_pos"]) == pos, h
        assert np.array_equal(n["mn"], lo) and np.array_equal(n["mx"], hi), h
    # unreachable slots keep the invalid marker
    mask = np.ones(len(b["nodes"]), bool)
    mask[reach] = False
    assert np.all(b["nodes"]["prim_count"][mask] == INVALID)


def test_build_partition_mode_matches_presorted(oracle, rtcore):
    xyz, idx = rtcore.gen_triangle_soup(20000, 0.004, seed=4)
    for depth in (10, 16, 20):
        a = oracle.build_scene(xyz, idx, depth, presorted=True)
        b = oracle.build_scene(xyz, idx, depth, presorted=False)
        ra, rb = oracle.reachable(a["nodes"]), oracle.reachable(b["nodes"])
        assert np.array_equal(ra, rb)
        na, nb = a["nodes"][ra], b["nodes"][ra]
        assert np.array_equal(na["prim_count"], nb["prim_count"])
        assert np.array_equal(na["prim_pos"], nb["prim_pos"])
        assert np.array_equal(na["mn"], nb["mn"]) and np.array_equal(na["mx"], nb["mx"])
        # presorted ids are the (code, id) order; the partition mode holds the
        # same multiset of ids in every leaf
        sc, si = oracle.sort_pairs(a["unsorted_codes"])
        assert np.array_equal(a["ids"], si) and np.array_equal(a["codes"], sc)
        for h in ra[(na["prim_count"] & 4) != 0]:
            pos, cnt = int(a["nodes"][h]["prim_pos"]), int(a["nodes"][h]["prim_count"]) >> 3
            assert sorted(a["ids"][pos:pos + cnt]) == sorted(b["ids"][pos:pos + cnt])


# ---------------------------------------------------------------------- trace
def _rays(o, d, tmax=1e16, self_id=-1):
    o = np.atleast_2d(np.asarray(o, np.float32))
    d = np.atleast_2d(np.asarray(d, np.float32))
    r = np.zeros(max(len(o), len(d)), RAY_DTYPE)
    r["o"], r["d"], r["refl"], r["tmax"], r["tri"] = o, d, 1.0, tmax, self_id
    return r


def mt_all(xyz, idx, o, d):
    """Moller-Trumbore of one ray against every triangle, numpy float32, the
    operation order of Intersection.hlsl:47-64."""
    f = np.float32
    A, B, Cc = xyz[idx[:, 0]], xyz[idx[:, 1]], xyz[idx[:, 2]]
    e1, e2 = (B - A).astype(f), (Cc - A).astype(f)
    o, d = o.astype(f), d.astype(f)

    def cross(a, b):
        return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                         a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                         a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1).astype(f)

    def dot(a, b):
        return ((a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]).astype(f)

    with np.errstate(all="ignore"):
        p = cross(np.broadcast_to(d, e2.shape), e2)
        det = (f(1.0) / dot(e1, p)).astype(f)
        tv = (o - A).astype(f)
        u = (dot(tv, p) * det).astype(f)
        q = cross(tv, e1)
        v = (dot(np.broadcast_to(d, q.shape), q) * det).astype(f)
        t = (dot(e2, q) * det).astype(f)
        ok = (u >= 0) & (v >= 0) & ((u + v) <= 1) & (t >= 0)
    return u, v, t, ok


def test_trace_single_triangle_analytic(oracle):
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    b = oracle.build_scene(xyz, idx, 4)
    rays = _rays([[0.25, 0.5, 2.0], [0.9, 0.9, 2.0], [0.25, 0.5, -1.0]], [[0, 0, -1], [0, 0, -1], [0, 0, -1]])
    h = oracle.trace(b["nodes"], b["ids"], xyz, idx, rays)
    assert h[0]["tri"] == 0 and h[0]["t"] == 2.0 and h[0]["u"] == 0.25 and h[0]["v"] == 0.5
    assert h[0]["visited"] == 1 + 65536 and h[0]["root"] == 1
    assert h[1]["tri"] == -1 and h[1]["u"] == -1 and h[1]["t"] == np.float32(1e16) and h[1]["root"] == 1
    assert h[2]["tri"] == -1  # behind the origin: box T1 < 0, one box test
    assert h[2]["visited"] == 1
    # the self id is skipped
    h = oracle.trace(b["nodes"], b["ids"], xyz, idx, _rays([0.25, 0.5, 2.0], [0, 0, -1], self_id=0))
    assert h[0]["tri"] == -1


def test_trace_visit_counter_two_leaves(oracle):
    # two triangles far apart on x: root + 2 leaves; a ray hitting the first
    # leaf visits root, leaf A (1 tri), leaf B box (culled: beyond best t)
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [10, 0, -5], [11, 0, -5], [10, 1, -5]], np.float32)
    idx = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    b = oracle.build_scene(xyz, idx, 6)
    assert oracle.reachable(b["nodes"]).tolist() == [1, 2, 3]
    h = oracle.trace(b["nodes"], b["ids"], xyz, idx, _rays([0.2, 0.2, 5.0], [0, 0, -1]))
    assert h[0]["tri"] == 0 and h[0]["t"] == 5.0
    assert h[0]["visited"] == 3 + 65536 and h[0]["root"] == 2


@pytest.mark.parametrize("scene", ["torus", "soup"])
def test_trace_brute_force(oracle, rtcore, scene):
    if scene == "torus":
        xyz, idx = rtcore.gen_torus(seed=2)
        cam = rtcore.camera_look_at((0, 0.5, 3), (0, 0, 0), (0, 1, 0), 60.0, 1.0)
        rays = oracle.raygen_pinhole(cam.as_array(), 64, 64)
    else:
        xyz, idx = rtcore.gen_triangle_soup(3000, 0.02, seed=7)
        rng = np.random.default_rng(5)
        o = rng.random((3000, 3)).astype(np.float32)
        d = rng.normal(size=(3000, 3)).astype(np.float32)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = _rays(o, d.astype(np.float32))
    b = oracle.build_scene(xyz, idx, 20)
    h = oracle.trace(b["nodes"], b["ids"], xyz, idx, rays)
    bad = 0
    for i, r in enumerate(rays):
        u, v, t, ok = mt_all(xyz, idx, r["o"], r["d"])
        if not ok.any():
            bad += h[i]["tri"] != -1
            continue
        tmin = t[ok].min()
        best = np.flatnonzero(ok & (t == tmin))
        if h[i]["tri"] not in best or h[i]["t"] != tmin:
            bad += 1
            continue
        j = h[i]["tri"]
        assert h[i]["u"] == u[j] and h[i]["v"] == v[j]
    assert bad <= max(1, len(rays) // 10000), bad
    assert (h["tri"] >= 0).sum() > len(rays) // 20


def test_trace_shadow_start_leaf_same_occlusion(oracle, rtcore):
    xyz, idx = rtcore.gen_triangle_soup(20000, 0.01, seed=8)
    b = oracle.build_scene(xyz, idx, 20)
    cam = rtcore.camera_look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 60.0, 1.0)
    rays = oracle.raygen_pinhole(cam.as_array(), 128, 128)
    h = oracle.trace(b["nodes"], b["ids"], xyz, idx, rays)
    hit = h["tri"] >= 0
    p = rays["o"][hit] + h["t"][hit][:, None] * rays["d"][hit]
    L = np.array([0.5, 2.0, -1.0], np.float32)
    ld = (L - p).astype(np.float32)
    dist = np.linalg.norm(ld, axis=1).astype(np.float32)
    sr = _rays(p, (ld / dist[:, None]).astype(np.float32), tmax=dist, self_id=h["tri"][hit])
    from_leaf = oracle.trace(b["nodes"], b["ids"], xyz, idx, sr, start=h["root"][hit])
    from_root = oracle.trace(b["nodes"], b["ids"], xyz, idx, sr)
    anyhit = oracle.trace(b["nodes"], b["ids"], xyz, idx, sr, anyhit=True)
    occ_leaf, occ_root, occ_any = from_leaf["tri"] >= 0, from_root["tri"] >= 0, anyhit["tri"] >= 0
    assert np.array_equal(occ_root, occ_any)
    assert (occ_leaf == occ_root).mean() >= 0.9999
    assert 0.05 < occ_root.mean() < 0.95
    # the closest occluder from the root equals the any-hit t only when found first
    assert np.all(anyhit["t"][occ_any] >= from_root["t"][occ_any])


def test_trace_from_any_start_node_finds_the_closest_hit(oracle, rtcore):
    """From any start node the trail loop visits every leaf (the start's
    subtree, then the sibling subtree at each level up to the root's
    children), so the closest t equals brute force from every start: the hit
    leaf of a first trace, the root, an internal node and a leaf's sibling.
    Counters: a start at node s tests level(s) boxes before any descent (s and
    the sibling at each level below the root), so from the root they equal
    the plain trace's and from elsewhere they differ."""
    xyz, idx = rtcore.gen_triangle_soup(4000, 0.02, seed=11)
    b = oracle.build_scene(xyz, idx, 16)
    rng = np.random.default_rng(4)
    n = 1500
    o = rng.uniform(-0.1, 1.1, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays = _rays(o, d)
    brute = oracle.trace_brute(xyz, idx, rays)
    first = oracle.trace(b["nodes"], b["ids"], xyz, idx, rays)
    assert (brute["tri"] >= 0).mean() > 0.15
    leaf = (b["nodes"]["prim_count"] & 4) != 0
    inner = np.flatnonzero(~leaf[2:64] & (b["nodes"]["prim_count"][2:64] > 0)) + 2
    starts = {"hit leaf": np.where(first["root"] > 1, first["root"], 1),
              "root": np.ones(n, np.int32),
              "internal": np.full(n, inner[0], np.int32),
              "sibling of hit leaf": np.where(first["root"] > 1, first["root"] ^ 1, 1)}
    for name, st in starts.items():
        h = oracle.trace(b["nodes"], b["ids"], xyz, idx, rays, start=st)
        assert np.array_equal(h["t"], brute["t"]), name
        same = h["tri"] == brute["tri"]
        assert same.mean() > 0.999, name  # the rest are exact ties in t
        if name == "root":
            assert np.array_equal(h["visited"], first["visited"])
    # boxes tested before any descent from a start at level L: L (s and L - 1 siblings)
    tiny = _rays(np.array([[5.0, 5.0, 5.0]]), np.array([[1.0, 0.0, 0.0]]))  # misses every box
    for s in (1, 2, 3, int(inner[0]), 37):
        h = oracle.trace(b["nodes"], b["ids"], xyz, idx, tiny, start=np.array([s], np.int32))
        assert (h["visited"][0] & 0xFFFF) == int(s).bit_length(), s


# ----------------------------------------------------------------- rays/shade
def test_raygen_numpy(oracle, rtcore):
    cam = rtcore.camera_look_at((0.3, 0.5, 3.0), (0.1, 0.0, -0.2), (0, 1, 0), 60.0, 640 / 360)
    c = cam.as_array()
    rays = oracle.raygen_pinhole(c, 640, 360)
    f = np.float32
    py, px = np.mgrid[0:360, 0:640]
    x = ((2 * px + 1).astype(f) - f(640)) / f(640)
    y = -(((2 * py + 1).astype(f) - f(360)) / f(360))
    sx, sy = (x * c[12]).astype(f), (y * c[13]).astype(f)
    d = np.stack([((c[3 + k] + sx * c[6 + k]).astype(f) + sy * c[9 + k]).astype(f) for k in range(3)], -1)
    ln = np.sqrt(((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]).astype(f)).astype(f)
    d = (d / ln[..., None]).astype(f)
    assert np.array_equal(rays["d"].reshape(360, 640, 3), d)
    assert np.all(rays["o"] == c[:3]) and np.all(rays["tmax"] == f(1e16)) and np.all(rays["tri"] == -1)
    # the centre column looks along fwd
    assert np.allclose(rays["d"].reshape(360, 640, 3)[180, 320], c[3:6], atol=5e-3)


def test_render_direct_analytic_shadow(oracle):
    # floor quad z=0 over [-4,4]^2, occluder quad z=1 over [-0.5,0.5]^2, light at (0,0,3);
    # sample points stay off the quads' shared diagonals, where the coplanar
    # neighbour of the hit triangle reports t = 0 (a real self-shadowing case)
    floor = np.array([[-4, -4, 0], [4, -4, 0], [4, 4, 0], [-4, 4, 0]], np.float32)
    occl = np.array([[-0.5, -0.5, 1], [0.5, -0.5, 1], [0.5, 0.5, 1], [-0.5, 0.5, 1]], np.float32)
    xyz = np.concatenate([floor, occl])
    idx = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]], np.uint32)
    b = oracle.build_scene(xyz, idx, 8)
    pts = np.array([[0.0, 0.0], [0.3, -0.2], [2.0, 1.0], [-3.0, 1.0]], np.float32)
    o = np.concatenate([pts, np.full((4, 1), 5.0, np.float32)], 1)
    o[0, :2] = [0.05, 0.02]  # avoid the diagonal edge of the occluder
    rays = _rays(o, np.tile([0, 0, -1], (4, 1)))
    h = oracle.trace(b["nodes"], b["ids"], xyz, idx, rays)
    assert list(h["tri"][:2] >= 2) == [True, True]  # first two rays hit the occluder top
    # shade points on the floor: rays from below the occluder's shadow region
    o2 = np.array([[0.05, 0.02, 0.5], [0.3, -0.2, 0.5], [2.0, 1.0, 0.5], [-3.0, 1.0, 0.5]], np.float32)
    rays2 = _rays(o2, np.tile([0, 0, -1], (4, 1)))
    h2 = oracle.trace(b["nodes"], b["ids"], xyz, idx, rays2)
    assert np.all(h2["tri"] <= 1) and np.all(h2["tri"] >= 0)
    rgba, occ = oracle.render_direct(b["nodes"], b["ids"], xyz, idx, rays2, h2, [0, 0, 3],
                                     ambient=0.1, albedo=0.8)
    L = np.array([0, 0, 3.0])
    for i in range(4):
        p = o2[i].astype(np.float64) - [0, 0, 0.5]
        in_shadow = np.all(np.abs(p[:2] * 2 / 3) <= 0.5)  # segment p -> L crosses z = 1 at 2/3 p
        ld = (p - L) / np.linalg.norm(p - L)
        expect = (0.1 + (0.0 if in_shadow else -ld[2])) * 0.8
        assert occ[i, 0] == (1 if in_shadow else 0)
        assert abs(rgba[i, 0] - expect) < 1e-5


# ------------------------------------------------------------------------- GI
def _form_factor_parallel_rect(px, py, x0, x1, y0, y1, c):
    """Form factor from a differential area at (px, py, 0) facing +z to the
    rectangle [x0, x1] x [y0, y1] at height c (corner superposition of the
    differential-area-to-parallel-rectangle closed form)."""
    def f(a, b):
        A, B = a / c, b / c
        return (A / np.sqrt(1 + A * A) * np.arctan(B / np.sqrt(1 + A * A)) +
                B / np.sqrt(1 + B * B) * np.arctan(A / np.sqrt(1 + B * B))) / (2 * np.pi)

    def g(x, y):
        return np.sign(x) * np.sign(y) * f(np.abs(x), np.abs(y))

    return g(x1 - px, y1 - py) - g(x0 - px, y1 - py) - g(x1 - px, y0 - py) + g(x0 - px, y0 - py)


def _floor_and_ceiling():
    floor = [[-50, -50, 0], [50, -50, 0], [50, 50, 0], [-50, 50, 0]]
    ceil = [[-0.5, -0.5, 1], [0.5, -0.5, 1], [0.5, 0.5, 1], [-0.5, 0.5, 1]]
    xyz = np.array(floor + ceil, np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]], np.uint32)
    return xyz, idx


def test_gi_oracle_sky_and_zero_bounce(oracle, rtcore):
    xyz, idx = _floor_and_ceiling()
    b = oracle.build_scene(xyz, idx, 6)
    down = rtcore.camera_look_at((0, 0, 0.5), (0, 0, -3), (0, 1, 0), 30.0, 1.0)
    img, rays = oracle.render_gi(b["nodes"], b["ids"], xyz, idx, down.as_array(), 8, 8, 5, 0, 0.7, 1.5, 9)
    assert np.all(img == 0.0) and rays == 8 * 8 * 5  # floor hit, no bounce allowed
    side = rtcore.camera_look_at((0, 0, 0.5), (0, 100, 30), (0, 0, 1), 10.0, 1.0)  # above the horizon, past the ceiling
    img, rays = oracle.render_gi(b["nodes"], b["ids"], xyz, idx, side.as_array(), 8, 8, 4, 3, 0.7, 1.5, 9)
    assert np.all(img == 4 * 1.5) and rays == 8 * 8 * 4  # every primary ray escapes


def test_gi_oracle_form_factor(oracle, rtcore):
    """Floor under a 1 x 1 ceiling at height 1, one bounce: a floor point at p
    returns albedo * sky * (1 - F(p)), F the closed-form form factor.  Pins the
    cosine-weighted sampling, the face-forward normal and the path loop."""
    xyz, idx = _floor_and_ceiling()
    b = oracle.build_scene(xyz, idx, 6)
    cam = rtcore.camera_look_at((0.1, -0.05, 0.5), (0.1, -0.05, 0.0), (0, 1, 0), 90.0, 1.0)
    W = H = 12
    spp = 4096
    img, _ = oracle.render_gi(b["nodes"], b["ids"], xyz, idx, cam.as_array(), W, H, spp, 1, 0.7, 1.0,
                              31, nthreads=8)
    est = img / spp
    c = cam.as_array()
    py, px = np.mgrid[0:H, 0:W]
    x = ((2 * px + 1) - W) / W * c[12]
    y = -((2 * py + 1) - H) / H * c[13]
    d = c[3:6] + x[..., None] * c[6:9] + y[..., None] * c[9:12]
    t = -c[2] / d[..., 2]
    fx, fy = c[0] + t * d[..., 0], c[1] + t * d[..., 1]
    expect = 0.7 * (1 - _form_factor_parallel_rect(fx, fy, -0.5, 0.5, -0.5, 0.5, 1.0))
    sigma = 0.7 * np.sqrt(expect / 0.7 * (1 - expect / 0.7) / spp)
    z = (est - expect) / sigma
    # the pixel footprint varies F by far less than sigma at this resolution
    assert np.abs(z).max() < 5.0
    assert abs(z.mean()) < 5.0 / np.sqrt(W * H)
