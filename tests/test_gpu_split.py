"""The split backward of the batched step (sv_config.split_backward): the
backward runs as two launches over two view ranges and the first range's
record gather overlaps the second range's backward on a second stream. The
arithmetic is the same, so three steps leave bit-identical parameters and
moments with and without the split: graph and eager launches, persistent
and one-workgroup-per-tile rasters, the fused loss and the residual buffer
(SSIM on), ring views and views of uneven sizes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")


def _scene(sv, orc, sizes, n=30000):
    g = sv.scene_gaussians(0, n, 1.0)
    cams = []
    for k, (w, h) in enumerate(sizes):
        az = 2.0 * np.pi * k / len(sizes)
        cams.append(sv.scene_camera([3.0 * np.sin(az), 0.3 * (k % 3), 3.0 * np.cos(az)], [0, 0, 0], w, h, 60.0))
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    targets = [orc.render_and_grad(tg, c, orc.view_params(w, h, sh_deg=3))["color"] for c, (w, h) in zip(cams, sizes)]
    return g, cams, targets


def _run(sv, g, cams, targets, **cfg):
    with sv.Context(capacity=g["pos"].shape[0]) as c:
        c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3, lr_scale=1e-3,
                     lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1, pos_lr_schedule=0, **cfg)
        c.set_gaussians(g)
        c.set_views(cams)
        for v, t in enumerate(targets):
            c.set_target(v, t)
        views = list(range(len(cams)))
        for _ in range(3):
            c.step_views(views, sv.MODE_BATCHED)
        return c.get_gaussians(), c.get_moments(1), c.get_moments(2)


@pytest.mark.parametrize("sizes", [[(256, 192)] * 4, [(200, 136), (96, 72), (330, 250), (64, 48), (160, 160)]],
                         ids=["ring4", "uneven5"])
@pytest.mark.parametrize("graph,persistent,w_ssim", [(1, 0, 0.0), (0, 1, 0.0), (1, 1, 0.2)])
def test_split_backward_is_bit_identical(sv, orc, sizes, graph, persistent, w_ssim):
    g, cams, targets = _scene(sv, orc, sizes)
    ref = _run(sv, g, cams, targets, split_backward=0, use_graph=graph, persistent=persistent, w_ssim=w_ssim)
    got = _run(sv, g, cams, targets, split_backward=1, use_graph=graph, persistent=persistent, w_ssim=w_ssim)
    for a, b in zip(got, ref):
        for k in KEYS:
            assert np.array_equal(a[k], b[k]), k
    print(f"[split] {len(sizes)} views, graph {graph}, persistent {persistent}, ssim {w_ssim}: identical", flush=True)
