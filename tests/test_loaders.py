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


# ---------------------------------------------------------------- shading attributes
MTL = """# two materials
newmtl red
Ka 0.1 0.1 0.1 0.5
Kd 1 0 0
Ks 0.5 0.5 0.5
Ns 500
d 0.25
map_Kd red.png
newmtl blue
Kd 0 0 1
Tr 0.75
illum 1
Ni 1.5
"""


def test_obj_materials_and_file_normals(rtcore, tmp_path):
    """importMaterials known answers (model_obj.cpp:1119-1297): Kd / Ks with
    three numbers get w = 1, Ka with four keeps its w, Ns / 1000, d is alpha,
    Tr is 1 - alpha, illum 1 zeroes the specular colour; unset fields keep
    ModelOBJ's newmtl defaults.  Triangles take the material active at their
    face; file normals are kept as written, zero where a corner has none."""
    (tmp_path / "cube.mtl").write_text(MTL)
    path = tmp_path / "cube.obj"
    path.write_text(OBJ)
    m = rtcore.load_model_shading(str(path), raw=True)
    mats = m["materials"]
    assert len(mats) == 2
    red, blue = mats
    assert red["ambient"].tolist() == [F(0.1), F(0.1), F(0.1), 0.5]
    assert red["diffuse"].tolist() == [1, 0, 0, 1]
    assert red["specular"].tolist() == [0.5, 0.5, 0.5, 1]
    assert red["shininess"] == F(0.5) and red["alpha"] == F(0.25)
    assert blue["ambient"].tolist() == [F(0.2), F(0.2), F(0.2), 1]
    assert blue["diffuse"].tolist() == [0, 0, 1, 1]
    assert blue["specular"].tolist() == [0, 0, 0, 1]
    assert blue["alpha"] == F(0.25) and blue["shininess"] == 0
    assert mats["id"].tolist() == [0, 1]
    assert m["mat_ids"].tolist() == [0, 0, 0, 0, 1, 1]
    n = m["normals"]
    assert n[:4].tolist() == [[0, 0, -1]] * 4 and n[4:8].tolist() == [[0, 0, 1]] * 4
    assert not n[8:].any()
    # positions and indices are those of the positions-only loader
    xyz, idx = rtcore.load_model(str(path), raw=True)
    assert np.array_equal(m["xyz"], xyz) and np.array_equal(m["idx"], idx)


def test_obj_without_materials_gets_default_material_data(rtcore, tmp_path):
    """No MTL file (or an unreadable one): Model installs one MaterialData()
    (MaterialData.cpp:17-24) and every triangle uses it."""
    path = tmp_path / "cube.obj"
    path.write_text(OBJ)  # names cube.mtl, which does not exist here
    m = rtcore.load_model_shading(str(path), raw=True)
    (mat,) = m["materials"]
    assert mat["diffuse"].tolist() == [0.5, 0.5, 0.5, F(0.8)]
    assert mat["ambient"].tolist() == [0, 0, 0, 1]
    assert mat["specular"].tolist() == [F(0.8), F(0.8), F(0.8), 1]
    assert mat["alpha"] == 0 and mat["shininess"] == 0
    assert not m["mat_ids"].any()


def test_single_material_and_unknown_names(rtcore, tmp_path):
    """One material: every triangle gets it (Model.cpp:361-362); an unknown
    usemtl name falls back to material 0 (model_obj.cpp:1073-1079)."""
    (tmp_path / "one.mtl").write_text("newmtl only\nKd 0.3 0.6 0.9\n")
    (tmp_path / "two.mtl").write_text("newmtl a\nKd 1 1 1\nnewmtl b\nKd 0 0 0\n")
    body = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\n"
    (tmp_path / "one.obj").write_text("mtllib one.mtl\n" + body + "usemtl nope\nf 1 2 3\nf 2 4 3\n")
    (tmp_path / "two.obj").write_text(body + "usemtl b\nf 1 2 3\nusemtl nope\nf 2 4 3\nusemtl b\nf 1 2 4\n"
                                      "mtllib two.mtl\n")
    one = rtcore.load_model_shading(str(tmp_path / "one.obj"), raw=True)
    assert len(one["materials"]) == 1 and not one["mat_ids"].any()
    assert one["materials"][0]["diffuse"].tolist() == [F(0.3), F(0.6), F(0.9), 1]
    # the material table is read before the faces, wherever mtllib stands
    two = rtcore.load_model_shading(str(tmp_path / "two.obj"), raw=True)
    assert two["mat_ids"].tolist() == [1, 0, 1]


def test_generated_normals(rtcore, tmp_path):
    """An OBJ with no vn gets generateNormals (model_obj.cpp:541-616): the
    normalised sum of the unnormalised face normals.  Pinned by geometry: the
    vertices of a flat fan get the plane's normal exactly, and the apex of a
    symmetric square pyramid the axis; the base corners lie along the sum of
    their two faces' area-weighted normals."""
    (tmp_path / "fan.obj").write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    fan = rtcore.load_model_shading(str(tmp_path / "fan.obj"), raw=True)
    assert fan["normals"].tolist() == [[0, 0, 1]] * 4
    pyr = "v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nv 0 2 0\n" \
          "f 1 5 2\nf 2 5 3\nf 3 5 4\nf 4 5 1\n"
    (tmp_path / "pyr.obj").write_text(pyr)
    p = rtcore.load_model_shading(str(tmp_path / "pyr.obj"), raw=True)
    apex = p["normals"][p["idx"][0][1]]
    assert np.allclose(apex, [0, 1, 0], atol=1e-7) and abs(apex[0]) == 0 and abs(apex[2]) == 0
    # corner (-1, 0, -1): faces (-z) and (-x) sides, each of normal length 4 / |.|
    c = p["normals"][p["idx"][0][0]].astype(np.float64)
    want = np.array([-1.0, 1.0, -1.0]) / np.sqrt(3.0)  # (0, 2, -4) + (-4, 2, 0), normalised
    assert np.allclose(c, want, atol=1e-6)
    assert abs(np.linalg.norm(c) - 1) < 1e-6


def test_txt_normals_are_normalised(rtcore, tmp_path):
    """Model::LoadFile normalises ".txt" normals (Normalize: v / |v|)."""
    path = tmp_path / "n.txt"
    path.write_text("3 1\n0 0 0 0 3 4 0 0\n1 0 0 2 0 0 0 0\n0 1 0 1 1 1 0 0\n0 1 2\n")
    m = rtcore.load_model_shading(str(path), raw=True)
    n = m["normals"]
    assert n[0].tolist() == [0, F(0.6), F(0.8)] and n[1].tolist() == [1, 0, 0]
    assert np.allclose(n[2], np.full(3, 1 / np.sqrt(3)), atol=1e-7)
    assert len(m["materials"]) == 1 and m["materials"][0]["alpha"] == 0


def test_mtl_property_before_newmtl_is_an_error(rtcore, tmp_path):
    (tmp_path / "bad.mtl").write_text("Kd 1 1 1\nnewmtl x\n")
    (tmp_path / "bad.obj").write_text("mtllib bad.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    with pytest.raises(rtcore.RtError, match="before any newmtl"):
        rtcore.load_model_shading(str(tmp_path / "bad.obj"))
