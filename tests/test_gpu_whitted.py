"""Whitted frames on the GPU (rt_render_whitted) against the oracle's
transcription of the reference's ray-tracing colour stage
(oracle/whitted_oracle.c, pinned by tests/test_whitted_oracle.py).  Both run
the same IEEE single operations in the same order and visit nodes in the same
order, so images are expected bit-exact; the acceptance bar written beside
each check is >= 99.99 % identical first-pass hit ids, <= 0.1 % of pixels
beyond 1e-5 relative, and PSNR >= 50 dB."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = np.float32
SHADOWS, PHONG, GLOSS, ONE_PASS, SPLIT, QUADS = 1, 2, 4, 8, 16, 32
FORM_BITS = {"wavefront": 0, "one-pass": ONE_PASS, "split": SPLIT, "split-quads": SPLIT | QUADS}


def psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    peak = max(1.0, float(np.abs(b).max()))
    return float("inf") if mse == 0 else 10 * np.log10(peak * peak / mse)


def vertex_normals(xyz, idx):
    n = np.zeros_like(xyz, dtype=np.float64)
    a, b, c = (xyz[idx[:, k]].astype(np.float64) for k in range(3))
    f = np.cross(b - a, c - a)
    for k in range(3):
        np.add.at(n, idx[:, k], f)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    ln[ln == 0] = 1
    return (n / ln).astype(F)


def materials(rtcore):
    m = np.zeros(5, rtcore.MATERIAL_DTYPE)
    # matte, mirror, glass, half-mirror, default-like
    m[0] = ((0.8, 0.6, 0.4, 1.0), (0.1, 0.1, 0.1, 1.0), (0.2, 0.2, 0.2, 1.0), 0.0, 1.0, 0, 0)
    m[1] = ((0.3, 0.3, 0.3, 0.2), (0.0, 0.0, 0.0, 1.0), (0.9, 0.9, 0.95, 1.0), 0.5, 1.0, 1, 0)
    m[2] = ((0.95, 0.97, 0.9, 0.6), (0.02, 0.02, 0.02, 1.5), (0.8, 0.8, 0.8, 0.7), 0.9, 0.0, 2, 0)
    m[3] = ((0.2, 0.7, 0.3, 0.5), (0.05, 0.1, 0.05, 1.2), (0.6, 0.6, 0.6, 1.0), 0.3, 0.5, 3, 0)
    m[4] = ((0.5, 0.5, 0.5, 0.8), (0.0, 0.0, 0.0, 1.0), (0.8, 0.8, 0.8, 1.0), 0.0, 0.0, 4, 0)
    return m


@pytest.fixture(scope="module")
def room(ctx, oracle, rtcore):
    """A torus over a floor, inside a triangle soup: smooth normals for Phong,
    every material kind, rays that bounce, refract and leave."""
    tx, ti = rtcore.gen_torus(0.35, 0.12, 96, 48, 0.0, 7)
    sx, si = rtcore.gen_triangle_soup(4000, 0.03, 8)
    floor = np.array([[-2, -2, -0.6], [2, -2, -0.6], [2, 2, -0.6], [-2, 2, -0.6]], F)
    xyz = np.concatenate([tx, sx * F(2.0) - F(1.0), floor]).astype(F)
    idx = np.concatenate([ti, si + len(tx), np.array([[0, 1, 2], [0, 2, 3]]) + len(tx) + len(sx)]).astype(np.uint32)
    normals = vertex_normals(xyz, idx)
    rng = np.random.default_rng(9)
    mat_ids = np.concatenate([np.full(len(ti), 2), rng.integers(0, 5, len(si)), [1, 1]]).astype(np.uint32)
    mats = materials(rtcore)
    s = ctx.add_scene(xyz, idx, 18)
    ctx.build(s)
    ctx.scene_shading(s, normals, mat_ids, mats)
    ref = oracle.build_scene(xyz, idx, 18)
    yield s, xyz, idx, normals, mat_ids, mats, ref
    ctx.scene_destroy(s)


def run_pair(ctx, oracle, rtcore, scene, xyz, idx, ref, w, h, cam_pos, light, normals=None, mat_ids=None,
             mats=None, **kw):
    cam = rtcore.camera_look_at(cam_pos, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 55.0, w / h)
    st = ctx.render_whitted(scene, cam, w, h, light, flags=rtcore.FLAG_KEEP_HITS, **kw)
    g, gh, gocc = ctx.frame_read(w, h, 1, hits=True)
    r, rh, rocc, nrays = oracle.render_whitted(ref["nodes"], ref["ids"], xyz, idx, cam.as_array(), w, h, light,
                                               normals=normals, mat_ids=mat_ids, materials=mats, nthreads=8, **kw)
    return st, g, gh, gocc[:, 0], r, rh, rocc, nrays


def check(tag, st, g, gh, gocc, r, rh, rocc, nrays):
    ids = (gh["tri"] == rh["tri"]).mean()
    close = np.isclose(g, r, rtol=1e-5, atol=1e-6).all(1)
    exact = (g == r).all(1).mean()
    p = psnr(g, r)
    print(f"{tag}: px={len(g)} ids={ids:.6f} exact={exact:.6f} close={close.mean():.6f} psnr={p:.1f} "
          f"rays gpu={st['primary_rays'] + st['shadow_rays'] + st['bounce_rays']} oracle={nrays}", flush=True)
    assert ids >= 0.9999, tag
    assert close.mean() >= 0.999, tag
    assert p >= 50.0, tag
    assert (gocc == rocc).mean() >= 0.9999, tag


CASES = [
    ("primary only, shadows", dict(bounces=0, switches=SHADOWS)),
    ("3 bounces, shadows + gloss", dict(bounces=3, switches=SHADOWS | GLOSS, env=(0.3, 0.4, 0.6))),
    ("5 bounces, Phong", dict(bounces=5, switches=SHADOWS | PHONG | GLOSS, env=(0.9, 0.8, 0.7), seed=77)),
    ("2 bounces, no shadows, white misses", dict(bounces=2, switches=PHONG, env_flag=0)),
    ("16 bounces", dict(bounces=16, switches=SHADOWS | PHONG, env=(1.0, 1.0, 1.0), frame=3)),
]


@pytest.mark.parametrize("form", ["wavefront", "one-pass", "split", "split-quads"])
@pytest.mark.parametrize("tag,kw", CASES, ids=[c[0] for c in CASES])
def test_whitted_room_matches_oracle(ctx, oracle, rtcore, room, tag, kw, form):
    """The four forms of the device frame (a compacted queue per bounce
    pass; one lane per pixel through every pass; pass 0 per pixel, then one
    lane per live path through the rest; the same with four lanes per live
    path) against the oracle."""
    s, xyz, idx, normals, mat_ids, mats, ref = room
    kw = dict(kw)
    kw["switches"] = kw["switches"] | FORM_BITS[form]
    out = run_pair(ctx, oracle, rtcore, s, xyz, idx, ref, 160, 120, (0.4, -1.7, 0.9), (0.8, -0.6, 1.5),
                   normals=normals, mat_ids=mat_ids, mats=mats, **kw)
    check(f"{tag} [{form}]", *out)
    st, nrays = out[0], out[-1]
    assert st["primary_rays"] + st["shadow_rays"] + st["bounce_rays"] == nrays


def test_whitted_forms_identical_at_1080p(ctx, rtcore, room):
    """The three device forms give identical frames, counters included, on a
    full-size image; then a smaller image reuses the larger queues."""
    s = room[0]
    for w, h in ((1920, 1080), (333, 197)):
        cam = rtcore.camera_look_at((0.4, -1.7, 0.9), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 55.0, w / h)
        kw = dict(bounces=4, env=(0.2, 0.3, 0.4), seed=3)
        out = {}
        for form, bit in FORM_BITS.items():
            st = ctx.render_whitted(s, cam, w, h, (0.8, -0.6, 1.5), switches=SHADOWS | PHONG | GLOSS | bit,
                                    flags=rtcore.FLAG_COUNT, **kw)
            out[form] = (st, ctx.frame_read(w, h, 1)[0])
        a, fa = out["wavefront"]
        print(f"{w}x{h}: " + ", ".join(f"{k} {v[0]['trace_ms']:.3f} ms" for k, v in out.items()) +
              f", rays {a['primary_rays'] + a['shadow_rays'] + a['bounce_rays']}", flush=True)
        for form in ("one-pass", "split", "split-quads"):
            b, fb = out[form]
            assert np.array_equal(fa, fb), form
            for k in ("primary_rays", "shadow_rays", "bounce_rays", "box_tests", "tri_tests"):
                assert a[k] == b[k], (form, k)


@pytest.mark.parametrize("bounces", [5, 12])
def test_whitted_groups_and_quads_identical_uncounted(ctx, rtcore, room, bounces):
    """Frames rendered without RT_FLAG_COUNT, where the one-pass form and the
    split form's remaining passes share each wave's live paths out to lane
    groups (up to 16 lanes a path once at most 32 remain) and the split-quads
    form gives every path four lanes, equal the wavefront form's one lane per
    ray, and so do their ray counts, at 1080p and on a small image."""
    s = room[0]
    for w, h in ((1920, 1080), (333, 197)):
        cam = rtcore.camera_look_at((0.4, -1.7, 0.9), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 55.0, w / h)
        kw = dict(bounces=bounces, env=(0.2, 0.3, 0.4), seed=11)
        out = {}
        for form in ("wavefront", "one-pass", "split", "split-quads"):
            st = ctx.render_whitted(s, cam, w, h, (0.8, -0.6, 1.5), switches=SHADOWS | PHONG | GLOSS | FORM_BITS[form],
                                    **kw)
            out[form] = (st, ctx.frame_read(w, h, 1)[0])
        a, fa = out["wavefront"]
        print(f"{w}x{h}, {bounces} bounces: " + ", ".join(f"{k} {v[0]['trace_ms']:.3f} ms" for k, v in out.items()),
              flush=True)
        for form in ("one-pass", "split", "split-quads"):
            b, fb = out[form]
            assert np.array_equal(fa, fb), form
            for k in ("primary_rays", "shadow_rays", "bounce_rays"):
                assert a[k] == b[k], (form, k)
    with pytest.raises(rtcore.RtError):  # quads belong to the split form
        ctx.render_whitted(s, cam, w, h, (0.8, -0.6, 1.5), switches=SHADOWS | QUADS, **kw)


def test_whitted_repeated_frames_identical(ctx, rtcore, room):
    """A repeated frame runs its tiles heaviest first (the last frame's tile
    costs): both forms give the same frame and counters every time."""
    s = room[0]
    w, h = 640, 360
    cam = rtcore.camera_look_at((0.4, -1.7, 0.9), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 55.0, w / h)
    kw = dict(bounces=3, env=(0.2, 0.3, 0.4), seed=5)
    for sw in (SHADOWS | PHONG, SHADOWS | PHONG | ONE_PASS, SHADOWS | PHONG | SPLIT, SHADOWS | PHONG | SPLIT | QUADS):
        ref = None
        for rep in range(3):
            st = ctx.render_whitted(s, cam, w, h, (0.8, -0.6, 1.5), switches=sw, flags=rtcore.FLAG_COUNT, **kw)
            f = ctx.frame_read(w, h, 1)[0]
            if ref is None:
                ref = (st, f)
                continue
            assert np.array_equal(ref[1], f), (sw, rep)
            for k in ("primary_rays", "shadow_rays", "bounce_rays", "box_tests", "tri_tests"):
                assert ref[0][k] == st[k], (sw, rep, k)


def test_whitted_default_material_and_flat(ctx, oracle, rtcore):
    """No rt_scene_shading: the default MaterialData (alpha 0: every hit
    refracts with index 1), flat normals."""
    xyz, idx = rtcore.gen_torus(0.4, 0.15, 64, 32, 1e-3, 3)
    s = ctx.add_scene(xyz, idx, 16)
    ctx.build(s)
    ref = oracle.build_scene(xyz, idx, 16)
    out = run_pair(ctx, oracle, rtcore, s, xyz, idx, ref, 96, 96, (0.2, -1.5, 0.8), (0.5, -0.5, 1.2),
                   bounces=4, switches=SHADOWS | GLOSS, env=(0.5, 0.6, 0.7))
    check("default material", *out)
    with pytest.raises(rtcore.RtError):
        ctx.render_whitted(s, rtcore.camera_look_at((0, -2, 0), (0, 0, 0), (0, 0, 1)), 8, 8, (0, 0, 1),
                           switches=PHONG)
    ctx.scene_destroy(s)


def test_whitted_batch_offsets(ctx, oracle, rtcore):
    """The second scene of a batch: normals, material ids and materials are
    addressed through that scene's vertex and triangle offsets."""
    a = rtcore.gen_triangle_soup(3000, 0.02, 11)
    b = rtcore.gen_torus(0.3, 0.1, 48, 24, 0.0, 12)
    first, second = ctx.batch_create([len(a[1]), len(b[1])], [len(a[0]), len(b[0])], [14, 14])
    rng = np.random.default_rng(13)
    mats = materials(rtcore)
    for sc, (xyz, idx) in zip((first, second), (a, b)):
        ctx.scene_upload(sc, xyz, idx)
    ctx.build(first)
    ids_b = rng.integers(0, 5, len(b[1])).astype(np.uint32)
    nb = vertex_normals(*b)
    ctx.scene_shading(first, vertex_normals(*a), rng.integers(0, 5, len(a[1])).astype(np.uint32), mats[::-1].copy())
    ctx.scene_shading(second, nb, ids_b, mats)
    ref = oracle.build_scene(b[0], b[1], 14)
    out = run_pair(ctx, oracle, rtcore, second, b[0], b[1], ref, 80, 64, (0.1, -1.2, 0.5), (0.3, -1.0, 1.0),
                   normals=nb, mat_ids=ids_b, mats=mats, bounces=3, switches=SHADOWS | PHONG)
    check("batch scene 1", *out)
    ctx.scene_destroy(first)
    ctx.scene_destroy(second)


def test_whitted_loaded_obj_with_mtl(ctx, oracle, rtcore, tmp_path):
    """An OBJ + MTL through rt_load_model_ex, uploaded with its normals,
    material ids and materials, rendered on the GPU and by the oracle."""
    (tmp_path / "m.mtl").write_text("newmtl gold\nKa 0.2 0.15 0.05\nKd 0.9 0.7 0.2\nKs 0.9 0.8 0.5\nd 0.6\n"
                                    "newmtl glass\nKa 0 0 0 1.45\nKd 0.95 0.95 1.0 0.3\nKs 0.5 0.5 0.5\nd 0.1\n")
    xyz, idx = rtcore.gen_torus(0.5, 0.2, 40, 20, 0.0, 5)
    lines = ["mtllib m.mtl"] + [f"v {x:.7g} {y:.7g} {z:.7g}" for x, y, z in xyz]
    for k, t in enumerate(idx):
        if k % 200 == 0:
            lines.append("usemtl " + ("gold", "glass")[(k // 200) % 2])
        lines.append("f " + " ".join(str(int(i) + 1) for i in t))
    (tmp_path / "m.obj").write_text("\n".join(lines) + "\n")
    m = rtcore.load_model_shading(str(tmp_path / "m.obj"))
    assert len(m["materials"]) == 2 and set(m["mat_ids"].tolist()) == {0, 1}
    s = ctx.add_scene(m["xyz"], m["idx"], 16)
    ctx.build(s)
    ctx.scene_shading(s, m["normals"], m["mat_ids"], m["materials"])
    ref = oracle.build_scene(m["xyz"], m["idx"], 16)
    out = run_pair(ctx, oracle, rtcore, s, m["xyz"], m["idx"], ref, 96, 72, (0.2, -1.3, 0.7), (0.6, -0.8, 1.0),
                   normals=m["normals"], mat_ids=m["mat_ids"], mats=m["materials"], bounces=4,
                   switches=SHADOWS | PHONG | GLOSS, env=(0.4, 0.5, 0.6))
    check("obj + mtl", *out)
    ctx.scene_destroy(s)


def test_scene_shading_rejects_bad_input(ctx, rtcore):
    xyz, idx = rtcore.gen_torus(0.4, 0.15, 16, 8, 0.0, 1)
    s = ctx.add_scene(xyz, idx, 10)
    with pytest.raises(rtcore.RtError):  # material id beyond the table
        ctx.scene_shading(s, None, np.full(len(idx), 5, np.uint32), materials(rtcore))
    with pytest.raises(rtcore.RtError):  # ids but no table: only material 0 exists
        ctx.scene_shading(s, None, np.ones(len(idx), np.uint32), None)
    ctx.scene_shading(s, None, np.zeros(len(idx), np.uint32), None)
    ctx.scene_destroy(s)


def test_autotune_whitted(ctx, rtcore, room, tmp_path):
    """The autotuner returns a valid shape and form, writes the cache, reads
    it back, and the chosen form renders the same frame as the other."""
    s = room[0]
    w, h = 320, 240
    cam = rtcore.camera_look_at((0.4, -1.7, 0.9), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 55.0, w / h)
    kw = dict(bounces=3, switches=SHADOWS | PHONG, env=(0.2, 0.3, 0.4))
    cache = tmp_path / "tune.txt"
    t1, sw1 = ctx.autotune_whitted(s, cam, w, h, (0.8, -0.6, 1.5), cache_path=str(cache), **kw)
    assert sw1 & ~(ONE_PASS | SPLIT | QUADS) == kw["switches"] and len(cache.read_text().splitlines()) == 1
    t2, sw2 = ctx.autotune_whitted(s, cam, w, h, (0.8, -0.6, 1.5), cache_path=str(cache), **kw)
    assert (t1, sw1) == (t2, sw2) and len(cache.read_text().splitlines()) == 1
    ctx.render_whitted(s, cam, w, h, (0.8, -0.6, 1.5), **dict(kw, switches=sw1))
    a = ctx.frame_read(w, h, 1)[0]
    for bit in FORM_BITS.values():
        ctx.render_whitted(s, cam, w, h, (0.8, -0.6, 1.5), **dict(kw, switches=kw["switches"] | bit))
        assert np.array_equal(a, ctx.frame_read(w, h, 1)[0]), bit
    with pytest.raises(rtcore.RtError):  # the two forms exclude each other
        ctx.render_whitted(s, cam, w, h, (0.8, -0.6, 1.5), **dict(kw, switches=kw["switches"] | ONE_PASS | SPLIT))
    ctx.set_tuning(128, 26, 8, 0)


def test_whitted_edge_cases(ctx, oracle, rtcore):
    """An empty scene (every ray misses), a 1 x 1 image, one bounce in the
    wavefront form, and a single-triangle scene on both structures."""
    empty = ctx.scene_create(16, 16, 8)
    ctx.scene_upload(empty, np.zeros((0, 3), F), np.zeros((0, 3), np.uint32))
    ctx.build(empty)
    cam = rtcore.camera_look_at((0, -2, 0.5), (0, 0, 0), (0, 0, 1), 50.0, 1.0)
    ctx.render_whitted(empty, cam, 17, 9, (0, 0, 2), bounces=3, env=(0.25, 0.5, 0.75))
    img = ctx.frame_read(17, 9, 1)[0]
    assert np.array_equal(img, np.tile(np.array([0.25, 0.5, 0.75, 1.0], F), (17 * 9, 1)))
    ctx.render_whitted(empty, cam, 5, 5, (0, 0, 2), bounces=2, env_flag=0)
    assert (ctx.frame_read(5, 5, 1)[0] == 1).all()
    ctx.scene_destroy(empty)
    xyz = np.array([[-1, 0, -1], [1, 0, -1], [0, 0, 1]], F)
    idx = np.array([[0, 1, 2]], np.uint32)
    one = ctx.add_scene(xyz, idx, 4)
    ctx.build(one)
    ref = oracle.build_scene(xyz, idx, 4)
    for w, h, bounces in ((1, 1, 0), (1, 1, 1), (7, 3, 1), (64, 48, 5)):
        cam = rtcore.camera_look_at((0.1, -2, 0.2), (0, 0, 0), (0, 0, 1), 50.0, w / h)
        for structure in (rtcore.STRUCTURE_LBVH, rtcore.STRUCTURE_BVH):
            if structure == rtcore.STRUCTURE_BVH:
                ctx.bvh_build(one)
            else:
                ctx.scene_select(one, rtcore.STRUCTURE_LBVH)
            ctx.render_whitted(one, cam, w, h, (0.5, -1.0, 1.0), bounces=bounces, switches=SHADOWS,
                               env=(0.1, 0.2, 0.3))
            g = ctx.frame_read(w, h, 1)[0]
            r = oracle.render_whitted(ref["nodes"], ref["ids"], xyz, idx, cam.as_array(), w, h, (0.5, -1.0, 1.0),
                                      bounces=bounces, switches=SHADOWS, env=(0.1, 0.2, 0.3))[0]
            assert np.array_equal(g, r), (w, h, bounces, structure)
    ctx.scene_destroy(one)
