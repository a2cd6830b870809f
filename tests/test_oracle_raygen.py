"""Pins of the oracle's jittered and camera-matrix ray generation and of the
direct-mode render job it composes from them -- CPU only.

  orc_pcg32_draws       the published pcg32-demo output of pcg-c-basic
                        (pcg32_srandom_r(42, 54): first six draws)
  orc_pcg_jitter        a pure-Python PCG32 (integer arithmetic, no numpy
                        wrap-around) and the unit mapping (x >> 8) / 2^24
  orc_raygen_pinhole_jitter
                        zero jitter is orc_raygen_pinhole bit for bit; a
                        float64 evaluation of the jittered pixel formula
  orc_raygen_world      closed forms: the identity camera gives
                        normalize(x, y, 2) with x = (2px + 1 - N)/N; a rotated
                        and translated camera gives the rotated directions
                        from the translation; a pinhole basis made of the
                        matrix columns (tan 1/2) gives the same rays
  render_direct_job     sky-only views add spp * sky; a lit floor seen from
                        above adds (ambient + cos) * albedo per sample, with
                        the hit point and cosine in float64
"""
import math

import numpy as np
import pytest

PCG_DEMO_42_54 = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
M64 = (1 << 64) - 1


def _pcg_py(initstate, initseq):
    st = [0, ((initseq << 1) | 1) & M64]

    def nxt():
        old = st[0]
        st[0] = (old * 6364136223846793005 + st[1]) & M64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    nxt()
    st[0] = (st[0] + initstate) & M64
    nxt()
    return nxt


def test_pcg32_published_vector(oracle):
    assert [int(x) for x in oracle.pcg32_draws(42, 54, 6)] == PCG_DEMO_42_54
    nxt = _pcg_py(42, 54)
    assert [nxt() for _ in range(6)] == PCG_DEMO_42_54


@pytest.mark.parametrize("seed,frame,sample", [(0, 0, 0), (5, 3, 1), (2**40 + 7, 65535, 9)])
def test_pcg_jitter_python(oracle, seed, frame, sample):
    w, h = 13, 7
    got = oracle.pcg_jitter(seed, frame, sample, w, h)
    init = (seed + 0x9E3779B97F4A7C15 * (frame * 65536 + sample + 1)) & M64
    for pix in range(w * h):
        nxt = _pcg_py(init, pix)
        u0 = (nxt() >> 8) / 16777216.0
        u1 = (nxt() >> 8) / 16777216.0
        # 2u - 1 is exact in float32 for a 24-bit u
        assert got[pix, 0] == np.float32(2 * u0 - 1) and got[pix, 1] == np.float32(2 * u1 - 1)
    big = oracle.pcg_jitter(seed, frame, sample, 256, 256)
    assert big.min() >= -1.0 and big.max() < 1.0 and abs(float(big.mean())) < 0.01


def test_raygen_jitter_zero_and_float64(oracle, rtcore):
    W, H = 96, 54
    cam = rtcore.camera_look_at((0.3, 0.5, 3.0), (0.1, 0.0, -0.2), (0, 1, 0), 60.0, W / H)
    c = cam.as_array()
    zero = np.zeros((W * H, 2), np.float32)
    assert oracle.raygen_pinhole(c, W, H, zero).tobytes() == oracle.raygen_pinhole(c, W, H).tobytes()
    jit = oracle.pcg_jitter(7, 1, 2, W, H)
    rays = oracle.raygen_pinhole(c, W, H, jit)
    cd = c.astype(np.float64)
    py, px = np.divmod(np.arange(W * H), W)
    x = ((2 * px + 1 + jit[:, 0].astype(np.float64)) - W) / W
    y = -(((2 * py + 1 + jit[:, 1].astype(np.float64)) - H) / H)
    d = cd[3:6] + (x * cd[12])[:, None] * cd[6:9] + (y * cd[13])[:, None] * cd[9:12]
    d /= np.linalg.norm(d, axis=1)[:, None]
    assert np.abs(rays["d"] - d).max() < 2e-6
    # jittered rays leave the pixel centre's ray by at most half a pixel pitch in angle
    ctr = oracle.raygen_pinhole(c, W, H)
    assert 0 < np.abs(rays["d"] - ctr["d"]).max() < 2.0 * c[12] / W


def _rot(axis, deg):
    a = np.asarray(axis, np.float64)
    a /= np.linalg.norm(a)
    t = math.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K


def _world(R, T):
    m = np.zeros((4, 4), np.float32)
    m[:3, :3] = R
    m[:3, 3] = T
    m[3, 3] = 1.0
    return m


def test_raygen_world_identity_closed_form(oracle):
    N = 64
    rays = oracle.raygen_world(_world(np.eye(3), (0, 0, 0)), N, N)
    py, px = np.divmod(np.arange(N * N), N)
    x = (2 * px + 1 - N) / N
    y = -(2 * py + 1 - N) / N
    d = np.stack([x, y, np.full_like(x, 2.0)], 1)
    d /= np.linalg.norm(d, axis=1)[:, None]
    assert np.all(rays["o"] == 0) and np.abs(rays["d"] - d).max() < 1e-6
    # top-left pixel looks left (-x) and up (+y); tmax and self as the reference's (CSPrimaryRays.hlsl:53-56)
    assert rays["d"][0, 0] < 0 and rays["d"][0, 1] > 0
    assert np.all(rays["tmax"] == np.float32(1e16)) and np.all(rays["tri"] == -1) and np.all(rays["tmin"] == 0)


def test_raygen_world_rotated_translated(oracle, rtcore):
    W, H = 80, 48  # wider than tall: the pixel pitch stays 2/H
    R = _rot((0.3, 1.0, -0.2), 37.0)
    T = np.array([0.25, -0.5, 1.75])
    m = _world(R, T)
    jit = oracle.pcg_jitter(3, 0, 0, W, H)
    rays = oracle.raygen_world(m, W, H, jit)
    assert np.all(rays["o"] == T.astype(np.float32))
    py, px = np.divmod(np.arange(W * H), W)
    x = ((2 * px + jit[:, 0]) + 1 - W) / H
    y = -((2 * py + jit[:, 1]) + 1 - H) / H
    local = np.stack([x, y, np.full_like(x, 2.0)], 1)
    d = local @ R.astype(np.float32).astype(np.float64).T
    d /= np.linalg.norm(d, axis=1)[:, None]
    assert np.abs(rays["d"] - d).max() < 2e-5
    # the same rays from a pinhole basis made of the matrix columns
    Rf = R.astype(np.float32)
    cam = rtcore.Camera()
    cam.pos[:] = list(T.astype(np.float32))
    cam.fwd[:], cam.right[:], cam.up[:] = list(Rf[:, 2]), list(Rf[:, 0]), list(Rf[:, 1])
    cam.tan_x, cam.tan_y = 0.5 * W / H, 0.5
    pin = oracle.raygen_pinhole(cam.as_array(), W, H, jit)
    assert np.abs(pin["d"] - rays["d"]).max() < 2e-5


def _floor_scene():
    xyz = np.array([[-50, -50, 0], [50, -50, 0], [50, 50, 0], [-50, 50, 0]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return xyz, idx


def test_direct_job_sky_only(oracle, rtcore):
    xyz, idx = _floor_scene()
    ref = oracle.build_scene(xyz, idx, 8)
    cam = rtcore.camera_look_at((0, 0, 5), (0, 0, 10), (0, 1, 0), 60.0, 1.0)  # looks away
    acc, shadow, nhit = oracle.render_direct_job(ref["nodes"], ref["ids"], xyz, idx, 16, 16, 3, [(0, 0, 4)],
                                                 sky=0.75, seed=4, cam14=cam.as_array())
    assert nhit == 0 and shadow == 0
    assert np.all(acc[..., :3] == np.float32(2.25)) and np.all(acc[..., 3] == 3)


def test_direct_job_lit_floor_closed_form(oracle, rtcore):
    xyz, idx = _floor_scene()
    ref = oracle.build_scene(xyz, idx, 8)
    W = H = 24
    spp, seed, frame = 3, 9, 2
    pos = np.array([0.2, -0.1, 3.0])
    cam = rtcore.camera_look_at(tuple(pos), (0.2, 0.0, 0.0), (0, 1, 0), 50.0, 1.0)
    L = np.array([0.5, 0.5, 2.0])
    ambient, albedo = 0.1, 0.8
    acc, shadow, nhit = oracle.render_direct_job(ref["nodes"], ref["ids"], xyz, idx, W, H, spp, [tuple(L)],
                                                 ambient, albedo, 0.0, seed, frame, cam14=cam.as_array())
    assert nhit == W * H * spp and shadow == W * H * spp
    want = np.zeros(W * H)
    c = cam.as_array().astype(np.float64)
    for s in range(spp):
        jit = oracle.pcg_jitter(seed, frame, s, W, H).astype(np.float64)
        py, px = np.divmod(np.arange(W * H), W)
        x = ((2 * px + 1 + jit[:, 0]) - W) / W
        y = -(((2 * py + 1 + jit[:, 1]) - H) / H)
        d = c[3:6] + (x * c[12])[:, None] * c[6:9] + (y * c[13])[:, None] * c[9:12]
        d /= np.linalg.norm(d, axis=1)[:, None]
        hit = c[:3] + (-c[2] / d[:, 2])[:, None] * d
        ld = L - hit
        cos = ld[:, 2] / np.linalg.norm(ld, axis=1)  # -n'.Ldir with n' = +z facing the camera
        want += (ambient + np.maximum(cos, 0.0)) * albedo
    assert np.abs(acc[..., 0].reshape(-1) - want).max() < 1e-5
    assert np.all(acc[..., 3] == spp)
