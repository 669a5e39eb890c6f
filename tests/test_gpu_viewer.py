"""The viewer path on the tile pipeline (sv_render_camera): a camera that
is not a training view renders exactly as the oracle's training forward
for that camera (image within 1e-5, alpha bit-exact), and rendering it
leaves the training views' images and the next training step unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_render_camera_matches_oracle_and_leaves_training_state(sv, orc):
    g = sv.scene_gaussians(31, 6000, 1.0)
    cams = sv.ring_cameras(2, 3.0, 96, 64)
    viewer = sv.scene_camera([1.1, 0.7, 2.6], [0.1, 0.0, 0.0], 160, 112, 55.0)
    view = orc.view_params(160, 112, sh_deg=3)
    ref = orc.render_and_grad(g, viewer, view)
    with sv.Context(capacity=6000) as c:
        c.set_config(sh_deg=3)
        c.set_gaussians(g)
        c.set_views(cams)
        c.forward(0)
        before = c.image(0)
        rgb, alpha = c.render_camera(viewer)
        after = c.image(0)
        c.forward(1)
        img1 = c.image(1)
    assert np.abs(rgb - ref["color"]).max() <= 1e-5
    assert np.array_equal(alpha.view(np.uint32), ref["alpha"].view(np.uint32))
    assert np.array_equal(before, after)
    want1 = orc.render_and_grad(g, cams[1], orc.view_params(96, 64, sh_deg=3))["color"]
    assert np.abs(img1 - want1).max() <= 1e-5
