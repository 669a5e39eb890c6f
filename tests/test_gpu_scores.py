"""Multi-view scoring on the device (sv_multiview_scores:
generateMultiViewScores, gaussian-renderer.ts:2754-2781, and
metric-map.wgsl:32-147) against the oracle: per-splat importance counts
bit-exact; pruning (counts x the view's mean l1, accumulated in float32)
within 1e-6 relative (the device sums the mean in another fixed order than
the host's sequential double sum); views without a target are skipped."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(sv, orc, n=8000, w=128, h=96, nv=3):
    g = sv.scene_gaussians(12, n, 1.0)
    cams = sv.ring_cameras(nv, 3.0, w, h)
    views = [orc.view_params(w, h, sh_deg=3) for _ in cams]
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(13, g["pos"], 0.03)
    targets = [orc.render_and_grad(tg, cam, v)["color"] for cam, v in zip(cams, views)]
    return g, cams, views, targets


def test_scores_match_oracle(sv, orc):
    g, cams, views, targets = _scene(sv, orc)
    with sv.Context(capacity=g["pos"].shape[0]) as c:
        c.set_config(sh_deg=3)
        c.set_gaussians(g)
        c.set_views(cams)
        for k, t in enumerate(targets):
            c.set_target(k, t)
        imp, pru, ns = c.multiview_scores([2, 0, 1])
    want_i, want_p, want_n = orc.multiview_scores(g, [cams[k] for k in (2, 0, 1)], [views[k] for k in (2, 0, 1)],
                                                  [targets[k] for k in (2, 0, 1)])
    print(f"[scores] {ns} views, {int((imp > 0).sum())} splats counted, max {int(imp.max())}", flush=True)
    assert ns == want_n == 3 and imp.max() > 0
    assert np.array_equal(imp, want_i)
    assert np.abs(pru.astype(np.float64) - want_p).max() <= 1e-6 * np.abs(want_p).max()


def test_scores_skip_views_without_target(sv, orc):
    g, cams, views, targets = _scene(sv, orc, nv=3)
    with sv.Context(capacity=g["pos"].shape[0]) as c:
        c.set_config(sh_deg=3)
        c.set_gaussians(g)
        c.set_views(cams)
        c.set_target(0, targets[0])
        c.set_target(1, targets[1])
        imp, pru, ns = c.multiview_scores([0, 1, 2])
    want_i, want_p, want_n = orc.multiview_scores(g, cams[:2], views[:2], targets[:2])
    assert ns == 2 and np.array_equal(imp, want_i)
    assert np.abs(pru.astype(np.float64) - want_p).max() <= 1e-6 * np.abs(want_p).max()
