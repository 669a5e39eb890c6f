"""sv_render_camera_noalloc (the viewer render through the fixed device
scratch, no allocation) against sv_render_camera, called through the C ABI:
two calls in a row, the second with fewer Gaussians and a smaller image, each
bit for bit equal to the allocating entry point; sizes past the scratch's
maxima are rejected with SV_E_CAPACITY (before running for Gaussians and
pixels, after projection for too many pairs), and a call after a rejected
one renders correctly again.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(sv, g):
    c = sv.Context(capacity=max(g["pos"].shape[0], 1))
    c.set_config(sh_deg=3)
    c.set_gaussians(g)
    return c


def _check(c, cam):
    want_rgb, want_a = c.render_camera(cam)
    got_rgb, got_a = c.render_camera_noalloc(cam)
    assert np.array_equal(got_rgb, want_rgb)
    assert np.array_equal(got_a, want_a)
    return float(np.abs(want_rgb).max())


def test_noalloc_render_matches_render_camera(sv):
    g1 = sv.scene_gaussians(0, 60000, 1.0)
    cam1 = sv.scene_camera([0.3, 0.4, 3.0], [0.0, 0.0, 0.0], 1024, 768, 60.0)
    c = _ctx(sv, g1)
    try:
        print("[noalloc] call 1: 60000 splats, 1024x768", flush=True)
        assert _check(c, cam1) > 0.0
        g2 = sv.scene_gaussians(7, 9000, 1.0)
        c.set_gaussians(g2)
        cam2 = sv.scene_camera([-2.0, 0.5, 2.0], [0.0, 0.0, 0.0], 320, 200, 50.0)
        print("[noalloc] call 2: 9000 splats, 320x200", flush=True)
        assert _check(c, cam2) > 0.0
    finally:
        c.close()


def test_noalloc_render_rejects_oversize(sv):
    g = sv.scene_gaussians(1, 70000, 1.0)
    c = _ctx(sv, g)
    try:
        cam = sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], 256, 256, 60.0)
        with pytest.raises(sv.SvError) as e:
            c.render_camera_noalloc(cam)
        assert e.value.status == sv.SV_E_CAPACITY
        small = {k: v[:30000] for k, v in g.items()}
        c.set_gaussians(small)
        big = sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], 2048, 1024, 60.0)
        with pytest.raises(sv.SvError) as e:
            c.render_camera_noalloc(big)
        assert e.value.status == sv.SV_E_CAPACITY
        # large splats: more than 2^20 (tile, splat) pairs, found after projection
        fat = dict(small)
        fat["logscale"] = (small["logscale"] + np.float32(3.0)).astype(np.float32)
        c.set_gaussians(fat)
        with pytest.raises(sv.SvError) as e:
            c.render_camera_noalloc(sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], 1024, 1024, 60.0))
        assert e.value.status == sv.SV_E_CAPACITY
        # the scratch is fine afterwards
        c.set_gaussians(small)
        print("[noalloc] after rejected calls: 30000 splats, 256x256", flush=True)
        assert _check(c, cam) > 0.0
    finally:
        c.close()
