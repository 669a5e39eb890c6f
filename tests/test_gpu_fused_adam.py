"""The batched step's fused Adam (inside the per-Gaussian backward) against
the same step with the gradients written first and Adam run over the
planes afterwards (keep_grads = 1): parameters and moments bit-identical
after two steps, so no part of the fused update reads a parameter that
another part has already moved."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")


def test_fused_adam_equals_separate_adam(sv):
    g = sv.scene_gaussians(41, 30_000, 1.0)
    cams = sv.ring_cameras(3, 3.0, 320, 240)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(42, g["pos"], 0.03)
    with sv.Context(capacity=30_000) as t:
        t.set_config(sh_deg=3)
        t.set_gaussians(tg)
        t.set_views(cams)
        targets = []
        for v in range(3):
            t.forward(v)
            targets.append(t.image(v))
    out = {}
    for keep in (0, 1):
        with sv.Context(capacity=30_000) as c:
            c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                         lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                         pos_lr_schedule=0, keep_grads=keep)
            c.set_gaussians(g)
            c.set_views(cams)
            for v in range(3):
                c.set_target(v, targets[v])
            for _ in range(2):
                c.step_views([0, 1, 2], sv.MODE_BATCHED)
            out[keep] = (c.get_gaussians(), c.get_moments(1), c.get_moments(2))
    for i in range(3):
        for k in KEYS:
            assert np.array_equal(out[0][i][k], out[1][i][k]), (i, k)
