"""Scene files (rt_load_model): the reference's .obj and .txt model formats
with its normalisation (ModelOBJ::normalize then Model::Scale).  The
normalised positions are checked against the same operations in numpy
float32, against invariants they must satisfy (centred box, farthest vertex
at 0.49), and the triangulation and vertex numbering against hand-derived
expectations."""
import numpy as np
import pytest

F = np.float32


def normalise_obj(p):
    """ModelOBJ::normalize(1, true) in float32, step by step."""
    lo, hi = p.min(0), p.max(0)
    centre = ((lo + hi) / F(2)).astype(F)
    ext = (hi - lo).astype(F)
    factor = F(1) / F(max(max(ext[0], ext[1]), ext[2]))
    return ((p + (-centre)).astype(F) * factor).astype(F)


def model_scale(p):
    """Model::Scale in float32, step by step."""
    mn, mx = p.min(0), p.max(0)
    mx = (mx - mn).astype(F)
    scale = F(1.8) / F(max(mx[0], max(mx[1], mx[2])))
    mn = (mn + (F(0.5) * mx).astype(F)).astype(F)
    p = (scale * (p - mn).astype(F)).astype(F)
    lng = ((p[:, 0] * p[:, 0] + p[:, 1] * p[:, 1]).astype(F) + p[:, 2] * p[:, 2]).astype(F)
    s = F(np.sqrt(F(lng.max())))
    s = F(F(0.49) / s)
    return (p * s).astype(F)


OBJ = """# a unit cube, quads, with normals and a negative index
mtllib cube.mtl
o cube
v 0 0 0
v 2 0 0
v 2 1 0
v 0 1 0
v 0 0 3
v 2 0 3
v 2 1 3
v 0 1 3
vn 0 0 -1
vn 0 0 1
vt 0 0
vt 1 0
usemtl red
f 1//1 2//1 3//1 4//1
f 5//2 8//2 7//2 6//2
usemtl blue
f 1/1 5/2 6/2
f -8 -7 -3
"""


def test_obj_triangulation_numbering_and_normalisation(rtcore, tmp_path):
    path = tmp_path / "cube.obj"
    path.write_text(OBJ)
    raw_xyz, raw_idx = rtcore.load_model(str(path), raw=True)
    # fans (0, k-1, k); a vertex per distinct (position index, uv, normal) in order of first use
    # the last face is 1 2 6 (-k = k-th most recent position); its corner 1 has
    # no uv and no normal, like corner 1 of the face before (uv (0, 0)), so it
    # reuses that vertex, while its corners 2 and 6 are new combinations
    assert raw_idx.tolist() == [[0, 1, 2], [0, 2, 3],
                                [4, 5, 6], [4, 6, 7],
                                [8, 9, 10],
                                [8, 11, 12]]
    pos = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0],
                    [0, 0, 3], [2, 0, 3], [2, 1, 3], [0, 1, 3]], F)
    expect = pos[[0, 1, 2, 3, 4, 7, 6, 5, 0, 4, 5, 1, 5]]
    assert np.array_equal(raw_xyz, expect)
    xyz, idx = rtcore.load_model(str(path))
    assert np.array_equal(idx, raw_idx)
    want = model_scale(normalise_obj(raw_xyz))
    assert np.array_equal(xyz, want)
    # invariants of the normalisation: centred box, farthest vertex at 0.49
    assert np.allclose((xyz.min(0) + xyz.max(0)) / 2, 0, atol=1e-6)
    assert abs(np.sqrt((xyz.astype(np.float64) ** 2).sum(1)).max() - 0.49) < 1e-6


def test_obj_faces_of_every_corner_form(rtcore, tmp_path):
    txt = "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nvn 0 0 1\n" \
          "f 1 2 3\nf 1/1 2/2 3/3\nf 1//1 2//1 3//1\nf 1/1/1 2/2/1 3/3/1\n"
    path = tmp_path / "forms.obj"
    path.write_text(txt)
    xyz, idx = rtcore.load_model(str(path), raw=True)
    assert len(idx) == 4
    # one vertex per distinct (position index, uv, normal) with absent values
    # zero: corner 1/1 (uv (0, 0)) is corner 1 again, and so is 1/1/1 of 1//1
    assert len(xyz) == 3 + 2 + 3 + 2
    assert idx[1][0] == idx[0][0] and idx[3][0] == idx[2][0]
    assert np.array_equal(xyz[idx[0]], xyz[idx[3]])


def test_txt_model(rtcore, tmp_path):
    rng = np.random.default_rng(3)
    p = rng.uniform(-5, 9, (40, 3)).astype(F)
    tris = rng.integers(0, 40, (30, 3))
    lines = ["40 30"]
    lines += [" ".join(f"{v:.9g}" for v in list(q) + [0, 0, 1, 0.5, 0.5]) for q in p]
    lines += [" ".join(str(int(i)) for i in t) for t in tris]
    path = tmp_path / "m.txt"
    path.write_text("\n".join(lines) + "\n")
    raw, idx = rtcore.load_model(str(path), raw=True)
    assert np.array_equal(raw, p) and np.array_equal(idx, tris)
    xyz, _ = rtcore.load_model(str(path))
    assert np.array_equal(xyz, model_scale(p))


def test_load_errors(rtcore, tmp_path):
    with pytest.raises(rtcore.RtError):
        rtcore.load_model(str(tmp_path / "missing.obj"))
    with pytest.raises(rtcore.RtError):
        rtcore.load_model(str(tmp_path / "model.ply"))
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n")
    with pytest.raises(rtcore.RtError, match="bad face corner"):
        rtcore.load_model(str(bad))
    empty = tmp_path / "empty.obj"
    empty.write_text("# nothing\nv 0 0 0\n")
    with pytest.raises(rtcore.RtError, match="no triangles"):
        rtcore.load_model(str(empty))
    badtxt = tmp_path / "bad.txt"
    badtxt.write_text("3 1\n0 0 0 0 0 1 0 0\n1 0 0 0 0 1 0 0\n0 1 0 0 0 1 0 0\n0 1 3\n")
    with pytest.raises(rtcore.RtError, match="in range"):
        rtcore.load_model(str(badtxt))
