"""Per-step statistics on the device (importance-accum.wgsl:19-40 and
densify-accum.wgsl:19-46, fused into the per-Gaussian backward) against
the oracle. One step over three views: every view of the step sees the same
parameters (Adam runs once, after the last view), so the oracle applies
accumulate_importance and accumulate_grads per view, in view order, to the
projections and geometric gradients of the unchanged scene. Bars: the
importance statistics (decayed tile counts, visibility, max radius) and
the accumulation counts bit-exact; the two gradient-norm sums within 1e-4
of the largest (the float gradients' bar); the unused slot zero."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfg(ctx, **kw):
    c = ctx.config()
    c.sh_deg = 3
    c.w_l2, c.w_l1, c.w_ssim = 0.8, 0.2, 0.0
    c.lr_pos = c.lr_sh = c.lr_opacity = c.lr_scale = c.lr_rot = 1e-3
    c.beta1, c.beta2, c.eps = 0.9, 0.999, 1e-15
    c.sh_update_interval = 1
    c.pos_lr_schedule = 0
    for k, v in kw.items():
        setattr(c, k, v)
    ctx.set_config(c)


@pytest.fixture(scope="module")
def scene(sv, orc):
    n, w, h, nv = 6000, 160, 120, 3
    g = sv.scene_gaussians(31, n, 1.0)
    cams = sv.ring_cameras(nv, 3.0, w, h)
    views = [orc.view_params(w, h, sh_deg=3) for _ in cams]
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(32, g["pos"], 0.03)
    targets = [orc.render_and_grad(tg, cam, v)["color"] for cam, v in zip(cams, views)]
    return g, cams, views, targets


def _want(orc, scene, order, wavefront=False):
    g, cams, views, targets = scene
    n = g["pos"].shape[0]
    imp = np.zeros((n, 4), np.float32)
    acc = np.zeros((n, 4), np.float32)
    for k in order:
        r = orc.render_and_grad(g, cams[k], views[k], target=targets[k], w_l2=0.8, w_l1=0.2,
                                wavefront=wavefront, wf_flags=0)
        imp = orc.accumulate_importance(r["proj"], imp, 0.98)
        acc = orc.accumulate_grads(r["grads"]["geom"], acc)
    return imp, acc


def _check(imp, acc, want_imp, want_acc, tag):
    vis = int((want_imp[:, 1] > 0).sum())
    cnt = int((want_acc[:, 3] > 0).sum())
    print(f"[stats] {tag}: {vis} splats visible, {cnt} with gradients", flush=True)
    assert vis > 0 and cnt > 0
    assert np.array_equal(imp[:, :3], want_imp[:, :3]), tag
    assert np.array_equal(acc[:, 3], want_acc[:, 3]), tag
    assert not acc[:, 1].any()
    for col in (0, 2):
        err = np.abs(acc[:, col].astype(np.float64) - want_acc[:, col]).max()
        assert err <= 1e-4 * np.abs(want_acc[:, col]).max(), (tag, col, err)


@pytest.mark.parametrize("mode", ["batched", "sequential"])
def test_stats_adjoint_modes(sv, orc, scene, mode):
    g, cams, views, targets = scene
    order = [2, 0, 1]
    with sv.Context(capacity=g["pos"].shape[0]) as c:
        _cfg(c)
        c.set_gaussians(g)
        c.set_views(cams)
        for k, t in enumerate(targets):
            c.set_target(k, t)
        c.step_views(order, sv.MODE_BATCHED if mode == "batched" else sv.MODE_SEQUENTIAL)
        imp, acc = c.stats()
    _check(imp, acc, *_want(orc, scene, order), mode)


def test_stats_plain_mode(sv, orc, scene):
    """The reference method (fixed-point wavefront backward) accumulates
    from its own geometric gradients."""
    g, cams, views, targets = scene
    order = [1, 2]
    with sv.Context(capacity=g["pos"].shape[0]) as c:
        _cfg(c, plain_wf_flags=0)
        c.set_gaussians(g)
        c.set_views(cams)
        for k, t in enumerate(targets):
            c.set_target(k, t)
        c.step_views(order, sv.MODE_PLAIN)
        imp, acc = c.stats()
    _check(imp, acc, *_want(orc, scene, order, wavefront=True), "plain")


def test_stats_accumulate_over_steps(sv, orc, scene):
    """Two steps: the second step's statistics decay the first's. The
    importance chain depends on the tile counts only, which the first Adam
    step (lr 1e-3) moves for a few splats; compare against the oracle run on
    the parameters the library holds before each step."""
    g, cams, views, targets = scene
    order = [0, 1, 2]
    with sv.Context(capacity=g["pos"].shape[0]) as c:
        _cfg(c)
        c.set_gaussians(g)
        c.set_views(cams)
        for k, t in enumerate(targets):
            c.set_target(k, t)
        c.step_views(order, sv.MODE_BATCHED)
        g1 = c.get_gaussians()
        imp1, _ = c.stats()
        c.step_views(order, sv.MODE_BATCHED)
        imp2, _ = c.stats()
    want = imp1.copy()
    for k in order:
        r = orc.render_and_grad(g1, cams[k], views[k])
        want = orc.accumulate_importance(r["proj"], want, 0.98)
    assert np.array_equal(imp2[:, :3], want[:, :3])
