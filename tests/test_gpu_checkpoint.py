"""Checkpoint / resume: training 3 steps straight equals training 2 steps,
saving, loading into a fresh context and training 1 more step, bit for bit
(parameters, moments, statistics, step counter and configuration travel);
a corrupted file is rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")


def _ctx(sv, g, cams, targets):
    c = sv.Context(capacity=32768)
    c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                 lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=2,
                 pos_lr_schedule=1, use_graph=1)
    if g is not None:
        c.set_gaussians(g)
    c.set_views(cams)
    for k, t in enumerate(targets):
        c.set_target(k, t)
    return c


def test_checkpoint_resume_is_exact(sv, orc, tmp_path):
    g = sv.scene_gaussians(0, 16384, 1.0)
    cams = sv.ring_cameras(3, 3.0, 200, 150)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    targets = [orc.render_and_grad(tg, cam, orc.view_params(200, 150, sh_deg=3))["color"] for cam in cams]
    views = [0, 1, 2]
    a = _ctx(sv, g, cams, targets)
    for _ in range(3):
        a.step_views(views, sv.MODE_BATCHED)
    want, wm2, wimp = a.get_gaussians(), a.get_moments(2), a.stats()[0]
    a.close()
    b = _ctx(sv, g, cams, targets)
    for _ in range(2):
        b.step_views(views, sv.MODE_BATCHED)
    path = str(tmp_path / "state.svckpt")
    b.save_checkpoint(path)
    b.close()
    c = _ctx(sv, None, cams, targets)
    c.set_config(sh_update_interval=7)  # overwritten by the checkpoint's configuration
    c.load_checkpoint(path)
    assert c.count == 16384 and c.step == 2 and c.config().sh_update_interval == 2
    c.step_views(views, sv.MODE_BATCHED)
    got, gm2, gimp = c.get_gaussians(), c.get_moments(2), c.stats()[0]
    for k in KEYS:
        assert np.array_equal(want[k], got[k]) and np.array_equal(wm2[k], gm2[k]), k
    assert np.array_equal(wimp, gimp)
    raw = bytearray(open(path, "rb").read())
    raw[100] ^= 0x40
    bad = str(tmp_path / "bad.svckpt")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(sv.SvError, match="checksum"):
        c.load_checkpoint(bad)
    c.close()
