"""Edge cases on the device against the oracle: a scene entirely behind the
camera (no pairs: background image, zero alpha, no last index, zero
gradients, a batched step that leaves every parameter and moment as it
was), a single splat on ragged views smaller than a tile and not a
multiple of 16 pixels, and a splat count that is not a multiple of the
256-thread blocks."""
import numpy as np
import pytest

from conftest import close_rel

pytestmark = pytest.mark.gpu

NO_LAST = 0xFFFFFFFF


def _cfg(ctx):
    c = ctx.config()
    c.sh_deg = 3
    c.w_l2, c.w_l1, c.w_ssim = 0.8, 0.2, 0.0
    c.lr_pos = c.lr_sh = c.lr_opacity = c.lr_scale = c.lr_rot = 1e-3
    c.beta1, c.beta2, c.eps = 0.9, 0.999, 1e-15
    c.sh_update_interval = 1
    c.pos_lr_schedule = 0
    ctx.set_config(c)


def test_scene_behind_the_camera(sv, orc):
    n, w, h = 3000, 96, 64
    g = sv.scene_gaussians(51, n, 1.0)
    g["pos"] = (g["pos"] + np.float32([0.0, 0.0, 10.0])).astype(np.float16).astype(np.float32)  # behind (0,0,3)
    cams = [sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, -1.0], w, h, 60.0)] * 2
    target = np.full((h, w, 3), 0.25, np.float32)
    ref = orc.render_and_grad(g, cams[0], orc.view_params(w, h, sh_deg=3), target=target, w_l2=0.8, w_l1=0.2)
    assert ref["n_isect"] == 0
    with sv.Context(capacity=n) as c:
        _cfg(c)
        c.set_gaussians(g)
        c.set_views(cams)
        for v in range(2):
            c.set_target(v, target)
        assert c.build_tiles(0) == 0
        c.forward(0)
        assert np.array_equal(c.image(0), ref["color"]) and not c.alpha(0).any()
        assert (c.last_ids(0) == NO_LAST).all()
        c.zero_grad()
        c.backward(0)
        for k, a in c.grads().items():
            assert not np.asarray(a).any(), k
        before = c.get_gaussians()
        st = c.step_views([0, 1], sv.MODE_BATCHED, stats=True)
        assert st.n_isect == 0
        after, m1 = c.get_gaussians(), c.get_moments(1)
    for k in before:
        assert np.array_equal(before[k], after[k]), k
        assert not m1[k].any(), k


@pytest.mark.parametrize("size", [(9, 7), (17, 33), (40, 23)])
def test_single_splat_ragged_views(sv, orc, size):
    w, h = size
    g = sv.scene_gaussians(52, 1, 1.0)
    g["pos"] = np.zeros((1, 3), np.float32)
    g["logscale"] = np.full((1, 3), -2.5, np.float32)
    g["logit"] = np.full((1,), 2.0, np.float32)
    cam = sv.scene_camera([0.2, -0.1, 3.0], [0.0, 0.0, 0.0], w, h, 60.0)
    view = orc.view_params(w, h, sh_deg=3)
    target = np.full((h, w, 3), 0.1, np.float32)
    ref = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    assert ref["n_isect"] >= 1
    with sv.Context(capacity=1) as c:
        _cfg(c)
        c.set_gaussians(g)
        c.set_views([cam])
        c.set_target(0, target)
        assert c.build_tiles(0) == ref["n_isect"]
        c.forward(0)
        assert np.abs(c.image(0) - ref["color"]).max() <= 1e-5
        assert np.array_equal(c.alpha(0).view(np.uint32), ref["alpha"].view(np.uint32))
        assert np.array_equal(c.last_ids(0), ref["last_ids"])
        c.zero_grad()
        c.backward(0)
        gg = c.grads()
    for k in ("logit", "sh", "pos", "logscale", "quat"):
        assert close_rel(gg[k], ref["grads"][k], 1e-4) <= 1e-4, k


def test_splat_count_off_the_block_size(sv, orc):
    n, w, h = 1000, 80, 48  # 1000 = 3 x 256 + 232
    g = sv.scene_gaussians(53, n, 1.0)
    cam = sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], w, h, 60.0)
    view = orc.view_params(w, h, sh_deg=3)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(54, g["pos"], 0.02)
    target = orc.render_and_grad(tg, cam, view)["color"]
    ref = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    with sv.Context(capacity=n) as c:
        _cfg(c)
        c.set_gaussians(g)
        c.set_views([cam])
        c.set_target(0, target)
        c.forward(0)
        assert np.array_equal(c.last_ids(0), ref["last_ids"])
        assert np.abs(c.image(0) - ref["color"]).max() <= 1e-5
        c.zero_grad()
        c.backward(0)
        gg = c.grads()
    for k in ("logit", "sh", "pos", "logscale", "quat"):
        assert close_rel(gg[k], ref["grads"][k], 1e-4) <= 1e-4, k
