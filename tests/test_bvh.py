"""The SAH BVH (rt_bvh_build_host; bvh.cpp:64-349 of the reference) and the
oracle's BVH_IntersectP (BVH.hlsl:124-217).

The builder is pinned by what the structure must satisfy: exact boxes
(inner = union of children, leaf = union of its triangles' bounds),
pre-order layout (first child next, second child at prim_pos), leaves that
partition the triangles, one triangle per leaf for max_prims = 1, the depth
bound, and known answers for the split rules (an outlier isolated by the SAH
buckets, the median order of small nodes, SAH beating the median and
midpoint splits on clustered geometry).  The oracle traversal is pinned by
brute force: its closest hit equals that of every triangle tested in turn."""
import numpy as np
import pytest

F = np.float32


def tri_bounds(xyz, idx):
    p = xyz[idx]  # (nt, 3, 3)
    return p.min(1), p.max(1)


def check_structure(nodes, ids, xyz, idx, max_prims):
    nt = len(idx)
    assert sorted(ids.tolist()) == list(range(nt))
    lo, hi = tri_bounds(xyz, idx)
    leaves = 0
    next_pos = 0
    stack = [(0, 0)]
    max_level = 0
    visited = 0
    while stack:
        i, level = stack.pop()
        visited += 1
        n = nodes[i]
        if n["prim_count"] > 0:
            cnt, pos = int(n["prim_count"]), int(n["prim_pos"])
            assert pos == next_pos  # leaves in pre-order cover consecutive ranges
            next_pos += cnt
            t = ids[pos:pos + cnt]
            assert np.array_equal(n["mn"], lo[t].min(0)) and np.array_equal(n["mx"], hi[t].max(0))
            if max_prims == 1:
                assert cnt == 1
            leaves += 1
            max_level = max(max_level, level)
        else:
            axis = -int(n["prim_count"])
            assert axis in (0, 1, 2)
            c0, c1 = i + 1, int(n["prim_pos"])
            assert c1 > c0
            a, b = nodes[c0], nodes[c1]
            assert np.array_equal(n["mn"], np.minimum(a["mn"], b["mn"]))
            assert np.array_equal(n["mx"], np.maximum(a["mx"], b["mx"]))
            stack.append((c1, level + 1))
            stack.append((c0, level + 1))
    assert next_pos == nt and visited == len(nodes) and len(nodes) == 2 * leaves - 1
    return leaves, max_level


SCENES = {
    "torus": lambda r: r.gen_torus(0.8, 0.3, 64, 32, 1e-3, 1),
    "soup": lambda r: r.gen_triangle_soup(20000, 0.01, 2),
    "sphere": lambda r: r.gen_noisy_sphere(65, 64, 0.05, 3),
    "mix": lambda r: r.gen_object_mix(30000, 4),
}


@pytest.mark.parametrize("scene", list(SCENES))
@pytest.mark.parametrize("split", [0, 1, 2])
@pytest.mark.parametrize("max_prims", [1, 4])
def test_structure_invariants(rtcore, scene, split, max_prims):
    xyz, idx = SCENES[scene](rtcore)
    nodes, ids, info = rtcore.bvh_build_host(xyz, idx, max_prims, split)
    leaves, max_level = check_structure(nodes, ids, xyz, idx, max_prims)
    assert info["leaves"] == leaves and info["nodes"] == len(nodes)
    assert info["max_depth"] == max_level + 1 <= 27


def test_sah_isolates_an_outlier(rtcore):
    """Five triangles near x = 0 and one at x = 10: every bucket split puts
    the five left and the outlier right at equal cost; the first minimum
    (bucket 0) wins, so the root's second child is the outlier's leaf."""
    base = np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0]], F)
    tris = [base + F(0.01 * k) for k in range(5)] + [base + np.array([10, 0, 0], F)]
    xyz = np.concatenate(tris).astype(F)
    idx = np.arange(18, dtype=np.uint32).reshape(6, 3)
    nodes, ids, _ = rtcore.bvh_build_host(xyz, idx, 1, rtcore.BVH_SAH)
    root = nodes[0]
    assert -int(root["prim_count"]) == 0  # split on x
    right = nodes[int(root["prim_pos"])]
    assert right["prim_count"] == 1 and ids[int(right["prim_pos"])] == 5


def test_small_nodes_split_at_the_median(rtcore):
    """Four or fewer triangles split at the median centroid of the widest
    axis, so leaves come out in centroid order along x."""
    base = np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0]], F)
    order = [3, 0, 2, 1]
    xyz = np.concatenate([base + np.array([order[k], 0, 0], F) for k in range(4)]).astype(F)
    idx = np.arange(12, dtype=np.uint32).reshape(4, 3)
    nodes, ids, _ = rtcore.bvh_build_host(xyz, idx, 1, rtcore.BVH_SAH)
    assert ids.tolist() == [1, 3, 2, 0]  # triangles at x = 0, 1, 2, 3


def test_sah_beats_median_and_midpoint(rtcore):
    xyz, idx = rtcore.gen_object_mix(40000, 9)
    cost = {s: rtcore.bvh_build_host(xyz, idx, 1, s)[2]["sah_cost"] for s in (0, 1, 2)}
    assert cost[0] < cost[1] and cost[0] < cost[2], cost


@pytest.mark.parametrize("scene", ["torus", "soup"])
def test_oracle_bvh_traversal_matches_brute_force(rtcore, oracle, scene):
    xyz, idx = SCENES[scene](rtcore)
    if scene == "soup":
        xyz, idx = rtcore.gen_triangle_soup(3000, 0.02, 5)
    rng = np.random.default_rng(1)
    n = 3000
    rays = np.zeros(n, oracle.RAY_DTYPE)
    lo, hi = xyz.min(0), xyz.max(0)
    rays["o"] = (lo + (hi - lo) * rng.uniform(-0.2, 1.2, (n, 3))).astype(F)
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F)
    rays["tmax"] = 1e16
    rays["tri"] = -1
    rays["tri"][::7] = rng.integers(0, len(idx), len(rays["tri"][::7]))
    brute = oracle.trace_brute(xyz, idx, rays)
    assert (brute["tri"] >= 0).mean() > 0.1
    nodes, ids, _ = rtcore.bvh_build_host(xyz, idx, 1, 0)
    for ref in (True, False):
        h = oracle.trace_bvh(nodes, ids, xyz, idx, rays, ref_leaf_loop=ref)
        assert np.array_equal(h["t"], brute["t"])
        same = h["tri"] == brute["tri"]
        assert same.mean() > 0.999  # the rest are exact ties in t
        assert np.array_equal(h["u"][same], brute["u"][same]) and np.array_equal(h["v"][same], brute["v"][same])
        hit = h["tri"] >= 0
        assert (nodes["prim_count"][h["root"][hit]] > 0).all()  # root is the hit's leaf
        assert (h["root"][~hit] == 0).all()
    a = oracle.trace_bvh(nodes, ids, xyz, idx, rays, ref_leaf_loop=True)
    b = oracle.trace_bvh(nodes, ids, xyz, idx, rays, ref_leaf_loop=False)
    # one-primitive leaves: the reference's loop tests the same triangle twice
    assert np.array_equal(a["visited"] & 0xFFFF, b["visited"] & 0xFFFF)
    assert np.array_equal(a["visited"] >> 16, 2 * (b["visited"] >> 16))


def test_reference_leaf_loop_misses_triangles_beyond_the_first(rtcore, oracle):
    """With several triangles per leaf the reference's loop (BVH.hlsl:165-169)
    only ever tests the leaf's first one; testing each once matches brute
    force."""
    xyz, idx = rtcore.gen_triangle_soup(3000, 0.05, 6)
    nodes, ids, info = rtcore.bvh_build_host(xyz, idx, 4, 0)
    assert (nodes["prim_count"] > 1).any()
    rng = np.random.default_rng(2)
    n = 2000
    rays = np.zeros(n, oracle.RAY_DTYPE)
    rays["o"] = rng.uniform(-0.1, 1.1, (n, 3)).astype(F)
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F)
    rays["tmax"] = 1e16
    rays["tri"] = -1
    brute = oracle.trace_brute(xyz, idx, rays)
    once = oracle.trace_bvh(nodes, ids, xyz, idx, rays, ref_leaf_loop=False)
    ref = oracle.trace_bvh(nodes, ids, xyz, idx, rays, ref_leaf_loop=True)
    assert np.array_equal(once["t"], brute["t"])
    assert (ref["t"] != brute["t"]).any()


def test_build_host_rejects_bad_input(rtcore):
    xyz, idx = rtcore.gen_torus(0.8, 0.3, 8, 4, 0.0, 1)
    with pytest.raises(rtcore.RtError):
        rtcore.bvh_build_host(xyz, idx, 0, 0)
    with pytest.raises(rtcore.RtError):
        rtcore.bvh_build_host(xyz, idx, 1, 3)
    bad = idx.copy()
    bad[0, 0] = len(xyz)
    with pytest.raises(rtcore.RtError):
        rtcore.bvh_build_host(xyz, bad, 1, 0)
