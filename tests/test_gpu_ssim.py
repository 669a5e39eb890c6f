"""The SSIM branch of the loss on the device (k_loss, training-loss.wgsl:33-84)
at the reference's default weights (w_l2, w_l1, w_ssim) = (0, 0.8, 0.2)
(gaussian-renderer.ts:506-522), against the oracle (orc_loss, pinned by the
closed forms of test_oracle_ssim.py):

- the residual image equals the oracle's on the same rendered image
  (1e-6 abs; both follow the shader's float order, so it is expected exact);
- gradients of a training step with the SSIM term on are within 1e-4 of
  max |ref| of the oracle's (legacy backward over the SSIM residual), in
  sequential mode, and the batched step captured as a graph gives the same
  numbers bit for bit;
- a captured step stays correct after sv_set_views replaces the cameras by
  the same views in swapped order (same view count and tile total, but
  other sizes and pixel offsets per slot, which the loss launches take by
  value): graph replay = eager launch, bit for bit.
"""
import numpy as np
import pytest

from conftest import close_rel

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")
W_L2, W_L1, W_SSIM = 0.0, 0.8, 0.2


def _cfg(ctx, **kw):
    ctx.set_config(sh_deg=3, w_l2=W_L2, w_l1=W_L1, w_ssim=W_SSIM, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                   lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                   pos_lr_schedule=0, **kw)


def _scene(sv, orc, sizes):
    g = sv.scene_gaussians(0, 16384, 1.0)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    cams = []
    for k, (w, h) in enumerate(sizes):
        az = 2.0 * np.pi * k / len(sizes)
        cams.append(sv.scene_camera([3.0 * np.sin(az), 0.0, 3.0 * np.cos(az)], [0, 0, 0], w, h, 60.0))
    targets = [orc.render_and_grad(tg, c, orc.view_params(w, h, sh_deg=3))["color"] for c, (w, h) in zip(cams, sizes)]
    return g, cams, targets


def test_ssim_residual_matches_oracle(sv, orc):
    g, cams, targets = _scene(sv, orc, [(256, 200)])
    with sv.Context(capacity=32768) as c:
        _cfg(c)
        c.set_gaussians(g)
        c.set_views(cams)
        c.set_target(0, targets[0])
        c.forward(0)
        img = c.image(0)
        c.loss(0)
        res = c.residual(0)
    want = orc.loss(img, targets[0], W_L2, W_L1, W_SSIM)
    ssim_only = orc.loss(img, targets[0], 0.0, 0.0, 1.0)
    err = float(np.abs(res - want).max())
    print(f"[ssim] residual max abs err {err:.3e}, exact {np.mean(res == want):.6f}, "
          f"|ssim term| max {np.abs(ssim_only).max():.3e}", flush=True)
    assert np.abs(ssim_only).max() > 1e-3, "the SSIM term must matter in this scene"
    assert err <= 1e-6


def test_ssim_step_gradients(sv, orc):
    sizes = [(256, 200), (200, 256)]
    g, cams, targets = _scene(sv, orc, sizes)
    ref = orc.zero_grads(g["pos"].shape[0])
    for cam, (w, h), t in zip(cams, sizes, targets):
        r = orc.render_and_grad(g, cam, orc.view_params(w, h, sh_deg=3), target=t, w_l2=W_L2, w_l1=W_L1,
                                w_ssim=W_SSIM)
        for k in KEYS:
            ref[k] += r["grads"][k]
    out = {}
    for mode, graph in ((sv.MODE_SEQUENTIAL, 0), (sv.MODE_BATCHED, 1)):
        with sv.Context(capacity=32768) as c:
            _cfg(c, keep_grads=1, use_graph=graph)
            c.set_gaussians(g)
            c.set_views(cams)
            for v, t in enumerate(targets):
                c.set_target(v, t)
            c.step_views([0, 1], mode)
            out[mode] = (c.grads(), c.get_gaussians())
    gq, pq = out[sv.MODE_SEQUENTIAL]
    gb, pb = out[sv.MODE_BATCHED]
    for k in KEYS:
        e = close_rel(gq[k], ref[k], 1e-4)
        print(f"[ssim] grad {k}: normalised max err vs oracle {e:.3e}", flush=True)
        assert e <= 1e-4, k
        assert np.array_equal(gq[k], gb[k]) and np.array_equal(pq[k], pb[k]), k


def _swap_run(sv, g, cams, targets, graph):
    c = sv.Context(capacity=32768)
    _cfg(c, keep_grads=1, use_graph=graph)
    c.set_gaussians(g)
    c.set_views(cams)
    for v, t in enumerate(targets):
        c.set_target(v, t)
    c.step_views([0, 1], sv.MODE_BATCHED)
    first = c.get_gaussians()
    c.set_views([cams[1], cams[0]])
    c.set_target(0, targets[1])
    c.set_target(1, targets[0])
    c.step_views([0, 1], sv.MODE_BATCHED)
    out = (first, c.get_gaussians(), c.grads(), [c.residual(v) for v in range(2)])
    c.close()
    return out


def test_graph_after_views_swapped(sv, orc):
    # 16 x 12 and 12 x 16 tiles: same tile total, different slot sizes
    sizes = [(256, 192), (192, 256)]
    g, cams, targets = _scene(sv, orc, sizes)
    eager = _swap_run(sv, g, cams, targets, 0)
    graph = _swap_run(sv, g, cams, targets, 1)
    for k in KEYS:
        assert np.array_equal(eager[0][k], graph[0][k]), ("first step", k)
        assert np.array_equal(eager[1][k], graph[1][k]), ("after swap", k)
        assert np.array_equal(eager[2][k], graph[2][k]), ("grads after swap", k)
    for a, b in zip(eager[3], graph[3]):
        assert np.array_equal(a, b)
    print("[ssim] graph replay after swapped views = eager", flush=True)
