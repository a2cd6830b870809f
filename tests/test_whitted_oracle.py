"""Pins the Whitted-frame oracle (oracle/whitted_oracle.c) against closed
forms: scenes of axis-aligned planes, where every hit point, normal, light
vector, mirror and Snell direction has an analytic value computed here in
float64 without any BVH.  Each test targets one term of the colour stage
(CSComputeColorRT.hlsl:18-262) so that a dropped term, a flipped sign or an
inverted index of refraction fails: the diffuse and ambient terms, the
specular power, shadow rays, the mirror continuation, refraction with
eta = 1, a glass slab (exit parallel to entry, Snell's law), total internal
reflection, the environment-mapping switch and Phong normals.  The last test
bounds the effect of starting rays at the root instead of the previous hit's
leaf (the reference's order): only exact-t ties can differ."""
import numpy as np
import pytest

F = np.float32
SHADOWS, PHONG, GLOSS = 1, 2, 4


def material(kd, ka, ks, alpha, shininess=0.0):
    m = np.zeros((), dtype=[("diffuse", "<f4", 4), ("ambient", "<f4", 4), ("specular", "<f4", 4),
                            ("shininess", "<f4"), ("alpha", "<f4"), ("id", "<u4"), ("pad", "<u4")])
    m["diffuse"], m["ambient"], m["specular"] = kd, ka, ks
    m["alpha"], m["shininess"] = alpha, shininess
    return m


def quad(z, half, up=True, cx=0.0, cy=0.0):
    """Two triangles covering [cx-half, cx+half] x [cy-half, cy+half] at height z,
    face normal +z (up) or -z."""
    a = [cx - half, cy - half, z]
    b = [cx + half, cy - half, z]
    c = [cx + half, cy + half, z]
    d = [cx - half, cy + half, z]
    xyz = np.array([a, b, c, d], F)
    idx = np.array([[0, 1, 2], [0, 2, 3]] if up else [[0, 2, 1], [0, 3, 2]], np.uint32)
    return xyz, idx


def merge(*parts):
    xyz, idx, off = [], [], 0
    for p, i in parts:
        xyz.append(p)
        idx.append(i + off)
        off += len(p)
    return np.concatenate(xyz), np.concatenate(idx).astype(np.uint32)


def render(oracle, rtcore, xyz, idx, cam_pos, cam_at, light, w=24, h=24, mats=None, mat_ids=None,
           normals=None, depth=8, **kw):
    s = oracle.build_scene(xyz, idx, depth)
    cam = rtcore.camera_look_at(cam_pos, cam_at, up=(0.0, 0.0, 1.0) if abs(cam_pos[0]) + abs(cam_pos[1]) > 0
                                else (0.0, 1.0, 0.0), fov_y_deg=50.0, aspect=w / h)
    cam14 = cam.as_array()
    rays = oracle.raygen_pinhole(cam14, w, h)
    rgba, hits, occ, n = oracle.render_whitted(s["nodes"], s["ids"], xyz, idx, cam14, w, h, light,
                                               normals=normals, mat_ids=mat_ids, materials=mats, **kw)
    return rgba, hits, occ, rays


def plane_hit(o, d, z):
    t = (z - o[:, 2]) / d[:, 2]
    return t, o + t[:, None] * d


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def lit_colour(kd, ka, ks, D, S, gloss=1.0):
    """Colour-stage formula in float64 for lit pixels (all four components)."""
    kd, ka, ks = (np.asarray(x, np.float64) for x in (kd, ka, ks))
    ks = ks.copy()
    ks[3] *= gloss
    return (ka[None] + D[:, None] * kd[3]) * kd[None] + (S[:, None] * (1 - kd[3])) * ks[None]


def inside(p, half, cx=0.0, cy=0.0, margin=1e-3):
    return (np.abs(p[:, 0] - cx) < half - margin) & (np.abs(p[:, 1] - cy) < half - margin)


def off_diagonal(p, cx=0.0, cy=0.0, margin=1e-3):
    """Away from a quad's shared edge, where a continuation ray leaving one
    triangle may graze its coplanar neighbour (as in the reference)."""
    return np.abs((p[:, 0] - cx) - (p[:, 1] - cy)) > margin


CAM = ((0.3, -1.2, 1.6), (0.0, 0.1, 0.0))


def test_diffuse_ambient_and_misses(oracle, rtcore):
    """One bounce-free pass over a lit plane: (Ka + max(0, -n.L) Kd.w) Kd; a
    miss is factor * env with w = 1."""
    xyz, idx = quad(0.0, 0.6)
    kd, ka, ks = (0.9, 0.6, 0.3, 1.0), (0.1, 0.2, 0.05, 1.0), (0.7, 0.7, 0.7, 0.5)
    light = (0.5, -0.4, 1.5)
    env = (0.25, 0.5, 0.75)
    rgba, hits, occ, rays = render(oracle, rtcore, xyz, idx, *CAM, light, mats=material(kd, ka, ks, 1.0)[None],
                                   bounces=0, switches=SHADOWS | GLOSS, env=env)
    o, d = rays["o"].astype(np.float64), rays["d"].astype(np.float64)
    t, p = plane_hit(o, d, 0.0)
    on = inside(p, 0.6)
    off = ~inside(p, 0.6, margin=-1e-3) | (t < 0)
    assert on.sum() > 100 and off.sum() > 50
    L = unit(p - np.array(light))
    D = np.maximum(0, -L[:, 2])
    want = lit_colour(kd, ka, ks, D, np.zeros_like(D))
    assert np.allclose(rgba[on], want[on], rtol=2e-6, atol=2e-6)
    assert np.array_equal(rgba[off], np.tile(np.array(env + (1.0,), F), (off.sum(), 1)))
    assert (hits["tri"][on] >= 0).all() and (hits["tri"][off] == -1).all()
    assert (occ[on] == 0).all() and (occ[off] == 2).all()


def test_back_face_lit_from_behind(oracle, rtcore):
    """Viewer and light both below a plane whose face normal points up: the
    normal is flipped to face the ray (:124-131), the pixel is lit, and the
    diffuse term uses the flipped normal: max(0, -(-n).L) = L.z > 0."""
    xyz, idx = quad(0.0, 0.7)
    kd, ka, ks = (0.6, 0.7, 0.8, 1.0), (0.05, 0.05, 0.05, 1.0), (0, 0, 0, 1)
    light = (0.2, -0.3, -1.2)
    rgba, _, occ, rays = render(oracle, rtcore, xyz, idx, (0.3, -1.2, -1.6), (0.0, 0.1, 0.0), light,
                                mats=material(kd, ka, ks, 1.0)[None], bounces=0, switches=SHADOWS)
    o, d = rays["o"].astype(np.float64), rays["d"].astype(np.float64)
    _, p = plane_hit(o, d, 0.0)
    on = inside(p, 0.7)
    L = unit(p - np.array(light))
    D = np.maximum(0, L[:, 2])
    assert on.sum() > 100 and D[on].min() > 0.3
    want = lit_colour(kd, ka, ks, D, np.zeros_like(D))
    assert np.allclose(rgba[on], want[on], rtol=2e-6, atol=2e-6)
    assert (occ[on] == 0).all()


def test_specular_power_49(oracle, rtcore):
    """Kd.w = 0 leaves Ka Kd + max(0, -L.R)^49 Ks with R the mirror direction."""
    xyz, idx = quad(0.0, 0.8)
    kd, ka, ks = (0.2, 0.3, 0.4, 0.0), (0.5, 0.5, 0.5, 1.0), (1.0, 0.8, 0.6, 1.0)
    light = (-0.25, 0.9, 1.4)  # mirror image of the camera side: a highlight in view
    rgba, _, _, rays = render(oracle, rtcore, xyz, idx, *CAM, light, w=32, h=32,
                              mats=material(kd, ka, ks, 1.0)[None], bounces=0, switches=SHADOWS)
    o, d = rays["o"].astype(np.float64), rays["d"].astype(np.float64)
    _, p = plane_hit(o, d, 0.0)
    on = inside(p, 0.8)
    L = unit(p - np.array(light))
    R = d * np.array([1, 1, -1])
    S = np.maximum(0, -(L * R).sum(1)) ** 49
    assert S[on].max() > 0.5  # the highlight is in the image
    want = lit_colour(kd, ka, ks, np.maximum(0, -L[:, 2]), S, gloss=0.0)
    assert np.allclose(rgba[on], want[on], rtol=1e-4, atol=1e-6)


def test_shadow_rays(oracle, rtcore):
    """An occluder between the light and the plane: plane points whose
    segment to the light crosses it get only the ambient term."""
    plane = quad(0.0, 0.8)
    occl = (np.array([[-0.2, -0.25, 0.8], [0.3, -0.2, 0.8], [0.0, 0.3, 0.8]], F), np.array([[0, 1, 2]], np.uint32))
    xyz, idx = merge(plane, occl)
    kd, ka, ks = (0.9, 0.9, 0.9, 1.0), (0.1, 0.1, 0.1, 1.0), (0, 0, 0, 1.0)
    light = (0.05, 0.0, 1.5)
    rgba, hits, occ, rays = render(oracle, rtcore, xyz, idx, (0.2, -1.5, 2.5), (0.0, 0.0, 0.0), light, w=48,
                                   h=48, mats=material(kd, ka, ks, 1.0)[None], bounces=0, switches=SHADOWS)
    o, d = rays["o"].astype(np.float64), rays["d"].astype(np.float64)
    _, p = plane_hit(o, d, 0.0)
    plane_px = inside(p, 0.8) & (hits["tri"] < 2)
    # where the segment p -> light crosses z = 0.8, and whether inside the occluder
    lt = np.array(light)
    s = (0.8 - p[:, 2]) / (lt[2] - p[:, 2])
    q = p + s[:, None] * (lt - p)
    A, B, C = occl[0].astype(np.float64)

    def edge(a, b, x):
        return (b[0] - a[0]) * (x[:, 1] - a[1]) - (b[1] - a[1]) * (x[:, 0] - a[0])

    e = np.stack([edge(A, B, q), edge(B, C, q), edge(C, A, q)], 1)
    shadowed = (e > 1e-4).all(1)
    clear = (e < -1e-4).any(1)
    assert (plane_px & shadowed).sum() > 30 and (plane_px & clear).sum() > 300
    L = unit(p - lt)
    lit = lit_colour(kd, ka, ks, np.maximum(0, -L[:, 2]), np.zeros(len(p)))
    dark = lit_colour(kd, ka, ks, np.zeros(len(p)), np.zeros(len(p)))
    dark[:, 3] = (ka[3] + 1.0 * kd[3]) * kd[3] + 1.0 * (1 - kd[3]) * ks[3]  # unlit D = S = (0, 0, 0, 1)
    m = plane_px & clear
    assert np.allclose(rgba[m], lit[m], rtol=2e-6, atol=2e-6)
    m = plane_px & shadowed
    assert np.allclose(rgba[m], dark[m], rtol=2e-6, atol=2e-6)
    assert (occ[plane_px & shadowed] == 1).all() and (occ[plane_px & clear] == 0).all()


def test_mirror_bounce(oracle, rtcore):
    """alpha = 1 always reflects: the second pass leaves the scene, adding
    Ks.rgb * env, and the alpha channel counts the pass."""
    xyz, idx = quad(0.0, 0.7)
    kd, ka, ks = (0.4, 0.4, 0.4, 1.0), (0.0, 0.0, 0.0, 1.0), (0.5, 0.25, 1.0, 1.0)
    env = (0.1, 0.2, 0.3)
    light = (0.0, 0.0, 2.0)
    one, _, _, rays = render(oracle, rtcore, xyz, idx, *CAM, light, mats=material(kd, ka, ks, 1.0)[None],
                             bounces=0, switches=SHADOWS, env=env)
    two, _, _, _ = render(oracle, rtcore, xyz, idx, *CAM, light, mats=material(kd, ka, ks, 1.0)[None],
                          bounces=3, switches=SHADOWS, env=env)
    _, p = plane_hit(rays["o"].astype(np.float64), rays["d"].astype(np.float64), 0.0)
    on = inside(p, 0.7) & off_diagonal(p)
    add = np.array([ks[0] * env[0], ks[1] * env[1], ks[2] * env[2], 1.0])
    assert np.allclose(two[on], one[on].astype(np.float64) + add, rtol=2e-6, atol=2e-6)
    miss = ~inside(p, 0.7, margin=-1e-3)
    assert np.array_equal(two[miss], one[miss])


def test_refraction_with_unit_index_goes_straight(oracle, rtcore):
    """alpha = 0 and Ka.w = 1: refract(d, n, 1) = d, so the second pass hits
    the lower plane on the primary ray's line, weighted by the upper Kd."""
    top = quad(0.0, 0.5)
    bottom = quad(-0.5, 2.0)
    xyz, idx = merge(top, bottom)
    m0 = material((0.5, 0.25, 0.75, 1.0), (0.1, 0.1, 0.1, 1.0), (0, 0, 0, 1), 0.0)
    m1 = material((0.2, 0.9, 0.4, 1.0), (0.05, 0.05, 0.05, 1.0), (0, 0, 0, 1), 1.0)
    mats = np.stack([m0, m1])
    mat_ids = np.array([0, 0, 1, 1], np.uint32)
    light = (0.3, 0.2, 1.0)
    rgba, _, _, rays = render(oracle, rtcore, xyz, idx, *CAM, light, mats=mats, mat_ids=mat_ids, bounces=1,
                              switches=0)
    o, d = rays["o"].astype(np.float64), rays["d"].astype(np.float64)
    _, p0 = plane_hit(o, d, 0.0)
    _, p1 = plane_hit(o, d, -0.5)
    on = inside(p0, 0.5) & inside(p1, 2.0) & off_diagonal(p0)
    lt = np.array(light)
    D0 = np.maximum(0, -unit(p0 - lt)[:, 2])
    D1 = np.maximum(0, -unit(p1 - lt)[:, 2])
    c0 = lit_colour(m0["diffuse"], m0["ambient"], m0["specular"], D0, np.zeros(len(D0)))
    c1 = lit_colour(m1["diffuse"], m1["ambient"], m1["specular"], D1, np.zeros(len(D1)))
    f = np.append(m0["diffuse"][:3].astype(np.float64), 1.0)
    want = c0 + c1 * f[None]
    assert on.sum() > 100
    assert np.allclose(rgba[on], want[on], rtol=1e-5, atol=1e-6)


def snell(d, n, eta):
    """Refracted unit direction for unit d, unit n facing d, relative index eta."""
    c1 = -(d * n).sum(1)
    c2 = np.sqrt(1 - eta * eta * (1 - c1 * c1))
    return unit(eta * d + (eta * c1 - c2)[:, None] * n)


def test_glass_slab_snell(oracle, rtcore):
    """A slab of index 1.5 (top face up, bottom face down): entering uses
    1 / 1.5, leaving 1.5 (the flipped-normal rule), so the exit ray is
    parallel to the primary ray, displaced; the floor it reaches is lit by a
    light whose falloff makes the colour depend on where it lands."""
    top = quad(0.0, 0.6)
    bot = quad(-0.15, 0.6, up=False)
    floor = quad(-0.6, 3.0)
    xyz, idx = merge(top, bot, floor)
    glass = material((0.9, 0.8, 0.7, 1.0), (0.0, 0.0, 0.0, 1.5), (0, 0, 0, 1), 0.0)
    flo = material((1.0, 1.0, 1.0, 1.0), (0.0, 0.0, 0.0, 1.0), (0, 0, 0, 1), 1.0)
    mats = np.stack([glass, flo])
    mat_ids = np.array([0, 0, 0, 0, 1, 1], np.uint32)
    light = (0.1, 0.2, -0.3)  # between slab and floor: only the floor is lit
    rgba, _, _, rays = render(oracle, rtcore, xyz, idx, (0.5, -1.4, 1.3), (0.0, 0.0, -0.3), light, w=32, h=32,
                              mats=mats, mat_ids=mat_ids, bounces=2, switches=0)
    o, d = rays["o"].astype(np.float64), rays["d"].astype(np.float64)
    up = np.array([0.0, 0.0, 1.0])
    _, p0 = plane_hit(o, d, 0.0)
    d1 = snell(d, np.tile(up, (len(d), 1)), 1 / 1.5)
    _, p1 = plane_hit(p0, d1, -0.15)
    d2 = snell(d1, np.tile(up, (len(d), 1)), 1.5)  # normal flipped to face the ray: +z again
    assert np.allclose(d2, d, atol=1e-12)  # Snell through parallel faces
    _, p2 = plane_hit(p1, d2, -0.6)
    on = inside(p0, 0.6) & inside(p1, 0.6) & inside(p2, 3.0) & off_diagonal(p0) & off_diagonal(p1)
    assert on.sum() > 100
    lt = np.array(light)
    # glass faces: light on the far side of each face from the viewer -> unlit,
    # so only the alpha channel gets (Ka.w + 1 Kd.w) Kd.w from them
    unlit_w = (1.5 + 1.0 * 1.0) * 1.0
    D2 = np.maximum(0, -unit(p2 - lt)[:, 2])
    kd = glass["diffuse"][:3].astype(np.float64)
    want = np.zeros((len(p0), 4))
    want[:, 3] = 2 * unlit_w + (1.0 + D2) * 1.0
    want[:, :3] = (D2[:, None] * 1.0) * (kd * kd)[None]
    assert np.allclose(rgba[on], want[on], rtol=1e-5, atol=1e-6)
    # a wrong index on the exit face moves the floor point and changes D2
    d2_bad = snell(d1, np.tile(up, (len(d), 1)), 1.0)
    _, p2b = plane_hit(p1, d2_bad, -0.6)
    assert np.abs(np.maximum(0, -unit(p2b - lt)[:, 2]) - D2)[on].max() > 1e-3


def test_total_internal_reflection(oracle, rtcore):
    """Seen from below (flipped normal, eta = 1.5), rays steeper than the
    critical angle refract (factor *= Kd), shallower ones reflect and keep
    factor 1; both then leave the scene and add factor * env."""
    xyz, idx = quad(0.0, 3.0)
    glass = material((0.5, 0.25, 0.125, 1.0), (0.0, 0.0, 0.0, 1.5), (0, 0, 0, 1), 0.0)
    env = (1.0, 1.0, 1.0)
    light = (0.0, 0.0, 1.0)  # above: never lights the underside seen from below
    rgba, _, _, rays = render(oracle, rtcore, xyz, idx, (0.0, -0.9, -0.6), (0.0, 0.2, 0.0), light, w=32, h=32,
                              mats=glass[None], bounces=1, switches=0, env=env)
    o, d = rays["o"].astype(np.float64), rays["d"].astype(np.float64)
    _, p = plane_hit(o, d, 0.0)
    on = inside(p, 3.0) & (d[:, 2] > 0) & off_diagonal(p)
    crit = np.sqrt(1 - 1 / 1.5 ** 2)  # |d.z| below this is past the critical angle
    tir = on & (d[:, 2] < crit - 1e-4)
    thru = on & (d[:, 2] > crit + 1e-4)
    assert tir.sum() > 50 and thru.sum() > 50
    assert np.allclose(rgba[tir, :3], 1.0, atol=1e-6)
    assert np.allclose(rgba[thru, :3], [0.5, 0.25, 0.125], atol=1e-6)


def test_environment_flag_zero_paints_misses_white(oracle, rtcore):
    xyz, idx = quad(0.0, 0.6)
    m = material((0.9, 0.6, 0.3, 1.0), (0.1, 0.2, 0.05, 1.0), (0.5, 0.5, 0.5, 1), 1.0)
    a, hits, _, _ = render(oracle, rtcore, xyz, idx, *CAM, (0, 0, 2), mats=m[None], bounces=0, env_flag=0)
    miss = hits["tri"] < 0
    assert miss.sum() > 10 and (a[miss] == 1).all() and (a[~miss] != 1).any()
    # with a bounce every path ends in a miss, which overwrites the result
    b, _, _, _ = render(oracle, rtcore, xyz, idx, *CAM, (0, 0, 2), mats=m[None], bounces=1, env_flag=0)
    assert (b == 1).all()


def test_phong_normals(oracle, rtcore):
    """With every vertex normal equal to a unit m, the interpolated normal is
    m: the plane is shaded as if tilted (diffuse -m'.L, lit test on m)."""
    xyz, idx = quad(0.0, 0.7)
    mvec = unit(np.array([[0.3, -0.2, 1.0]]))[0]
    normals = np.tile(mvec.astype(F), (4, 1))
    kd, ka, ks = (0.8, 0.7, 0.6, 1.0), (0.1, 0.1, 0.1, 1.0), (0, 0, 0, 1)
    light = (0.4, 0.5, 1.2)
    flat, _, _, rays = render(oracle, rtcore, xyz, idx, *CAM, light, mats=material(kd, ka, ks, 1.0)[None],
                              bounces=0, switches=SHADOWS)
    ph, _, _, _ = render(oracle, rtcore, xyz, idx, *CAM, light, mats=material(kd, ka, ks, 1.0)[None],
                         bounces=0, switches=SHADOWS | PHONG, normals=normals)
    o, d = rays["o"].astype(np.float64), rays["d"].astype(np.float64)
    _, p = plane_hit(o, d, 0.0)
    on = inside(p, 0.7)
    L = unit(p - np.array(light))
    n = np.tile(mvec, (len(p), 1))
    flip = (n * d).sum(1) > 0
    m2 = np.where(flip[:, None], -n, n)
    D = np.maximum(0, -(m2 * L).sum(1))
    lit = ((n * L).sum(1) * (n * d).sum(1)) > 0
    want = lit_colour(kd, ka, ks, np.where(lit, D, 0), np.zeros(len(D)))
    sel = on & lit
    assert sel.sum() > 100
    assert np.allclose(ph[sel], want[sel], rtol=1e-5, atol=1e-6)
    assert np.abs(ph[sel] - flat[sel]).max() > 1e-2


@pytest.mark.parametrize("seed", [5])
def test_root_start_matches_reference_leaf_start(oracle, rtcore, seed):
    """Starting rays at the root (the device's order) instead of the previous
    hit's leaf (the reference's) changes only exact-t ties."""
    xyz, idx = rtcore.gen_triangle_soup(3000, 0.03, seed)
    rng = np.random.default_rng(seed)
    mats = np.stack([material((0.8, 0.6, 0.4, 0.7), (0.1, 0.1, 0.1, 1.5), (0.9, 0.9, 0.9, 1.0), a)
                     for a in (0.0, 0.5, 1.0)])
    mat_ids = rng.integers(0, 3, len(idx)).astype(np.uint32)
    kw = dict(mats=mats, mat_ids=mat_ids, bounces=3, switches=SHADOWS | GLOSS, env=(0.2, 0.3, 0.4), depth=14,
              w=40, h=40)
    a, ha, _, _ = render(oracle, rtcore, xyz, idx, (0.1, -1.6, 0.4), (0, 0, 0), (1.0, -1.0, 1.0), **kw)
    b, hb, _, _ = render(oracle, rtcore, xyz, idx, (0.1, -1.6, 0.4), (0, 0, 0), (1.0, -1.0, 1.0),
                         ref_order=True, **kw)
    assert np.array_equal(ha["tri"], hb["tri"])
    differ = (a != b).any(1).mean()
    assert differ <= 2e-3
