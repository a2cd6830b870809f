"""rt_present: the device conversion of a float frame to the 8-bit RGBA a
R8G8B8A8_UNORM back buffer holds, against the same rule in numpy float32
(NaN -> 0, clamp to [0, 1], x 255, + 0.5, floor), on a Whitted frame whose
environment colour is out of range or NaN, and on a GI accumulation buffer
averaged by its sample count."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def unorm8(x):
    x = np.asarray(x, np.float32)
    c = np.clip(np.nan_to_num(x, nan=0.0, posinf=1.0, neginf=0.0), 0.0, 1.0).astype(np.float32)
    return np.floor(c * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)


def test_present_frame_buffer(ctx, rtcore, tmp_path):
    xyz, idx = rtcore.gen_torus(0.4, 0.15, 48, 24, 0.0, 2)
    s = ctx.add_scene(xyz, idx, 14)
    ctx.build(s)
    w, h = 97, 61
    cam = rtcore.camera_look_at((0.2, -1.4, 0.6), (0, 0, 0), (0, 0, 1), 55.0, w / h)
    ctx.render_whitted(s, cam, w, h, (0.5, -0.5, 1.0), bounces=2, switches=rtcore.WHITTED_SHADOWS,
                       env=(2.0, -1.0, float("nan")))
    f = ctx.frame_read(w, h, 1)[0]
    img = ctx.present(w, h)
    want = unorm8(f).reshape(h, w, 4)
    assert np.isnan(f).any() and (f > 1).any() and (f < 0).any()
    assert np.array_equal(img, want)
    rtcore.image_write(str(tmp_path / "frame.png"), img)
    ctx.scene_destroy(s)


def test_present_accumulation_average(ctx, rtcore):
    xyz, idx = rtcore.gen_torus(0.4, 0.15, 48, 24, 0.0, 2)
    s = ctx.add_scene(xyz, idx, 14)
    ctx.build(s)
    w, h = 64, 48
    acc = ctx.alloc(w * h * 16)
    ctx.zero(acc, w * h * 16)
    cam = rtcore.camera_look_at((0.2, -1.4, 0.6), (0, 0, 0), (0, 0, 1), 55.0, w / h)
    ctx.render_jobs([rtcore.Context.job(s, cam, w, h, acc, spp=3, max_bounces=2, sky=1.3)], want_stats=False)
    a = ctx.read_accum(acc, w, h)
    img = ctx.present(w, h, d_rgba=acc, average=True)
    avg = (a[..., :3] / a[..., 3:4]).astype(np.float32)
    want = unorm8(np.concatenate([avg, np.ones_like(a[..., :1])], -1))
    assert np.array_equal(img, want)
    ctx.free(acc)
    ctx.scene_destroy(s)


def test_present_rejects_mismatched_size(ctx, rtcore):
    xyz, idx = rtcore.gen_torus(0.4, 0.15, 16, 8, 0.0, 2)
    s = ctx.add_scene(xyz, idx, 10)
    ctx.build(s)
    cam = rtcore.camera_look_at((0.2, -1.4, 0.6), (0, 0, 0), (0, 0, 1), 55.0, 1.0)
    ctx.render_whitted(s, cam, 1, 1, (0.5, -0.5, 1.0), bounces=0)
    assert ctx.present(1, 1).shape == (1, 1, 4)
    with pytest.raises(rtcore.RtError):
        ctx.present(2, 2)  # the frame buffer holds a 1 x 1 frame
    ctx.scene_destroy(s)
