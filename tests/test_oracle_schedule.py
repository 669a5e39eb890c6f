"""Pins of the oracle's step orchestration used by the trajectory tests:
the position learning-rate schedule (computePositionLearningRate,
gaussian-renderer.ts:284-298), the per-tensor Adam step with the SH update
interval (run_adam_updates, :2455-2545) and the multi-view step (trainStep,
:3638-3660).
"""
import math

import numpy as np


def test_position_lr_endpoints(orc):
    lr = 1.6e-4
    f = lambda x: float(np.float32(x))  # noqa: E731
    # t = 0: the delay factor is 0.01 and the log-lerp is at lr
    assert orc.position_lr(lr, 0) == f(0.01 * float(np.float32(lr)))
    # t = 3000 (end of the warm-up): delay 1, a tenth of the way in log space
    assert abs(orc.position_lr(lr, 3000) - math.exp(0.9 * math.log(lr) + 0.1 * math.log(1.6e-6))) <= 1e-6 * lr
    # t >= 30000: the final rate
    assert orc.position_lr(lr, 30000) == f(1.6e-6)
    assert orc.position_lr(lr, 50000) == f(1.6e-6)
    # the sine warm-up dominates the early steps (its slope vanishes at 3000)
    warm = [orc.position_lr(lr, t) for t in range(0, 1501, 250)]
    decay = [orc.position_lr(lr, t) for t in range(3000, 30001, 1000)]
    assert all(a < b for a, b in zip(warm, warm[1:]))
    assert all(a > b for a, b in zip(decay, decay[1:]))


def _state(n=5, seed=0):
    rng = np.random.default_rng(seed)
    w = {"pos": 3, "logscale": 3, "quat": 4, "logit": 1, "sh": 48}
    mk = lambda: {k: rng.normal(size=(n, v) if v > 1 else (n,)).astype(np.float32) for k, v in w.items()}  # noqa: E731
    return mk(), mk()


def test_sh_interval_and_zero_position_rate(orc):
    p, g = _state()
    m1 = {k: np.zeros_like(v) for k, v in p.items()}
    m2 = {k: np.zeros_like(v) for k, v in p.items()}
    p0 = {k: v.copy() for k, v in p.items()}
    lrs = {"pos": 0.0, "logscale": 1e-3, "quat": 1e-3, "logit": 1e-3, "sh": 1e-3}
    for t in range(1, 16):
        orc.adam_update(p, m1, m2, g, t, lrs, 0.9, 0.999, 1e-15, sh_update_interval=16)
    assert np.array_equal(p["sh"], p0["sh"]) and not m1["sh"].any()
    assert np.array_equal(p["pos"], p0["pos"]) and not m1["pos"].any()
    assert not np.array_equal(p["logit"], p0["logit"])
    orc.adam_update(p, m1, m2, g, 16, lrs, 0.9, 0.999, 1e-15, sh_update_interval=16)
    # the first SH update starts from zero moments but its bias terms use the
    # shared step counter (t = 16): -lr (0.1 / sqrt(0.001)) sqrt(1 - b2^16) / (1 - b1^16) sign(g)
    d = (p["sh"].astype(np.float64) - p0["sh"])
    k16 = 0.1 / math.sqrt(0.001) * math.sqrt(1 - 0.999 ** 16) / (1 - 0.9 ** 16)
    assert np.allclose(d, -1e-3 * k16 * np.sign(g["sh"]), atol=2e-7)


def test_train_one_step_is_render_plus_adam(sv, orc):
    g = sv.scene_gaussians(0, 2048, 1.0)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    cam = sv.scene_camera([0.0, 0.0, 3.0], [0, 0, 0], 64, 64, 60.0)
    view = orc.view_params(64, 64, sh_deg=3)
    target = orc.render_and_grad(tg, cam, view)["color"]
    lrs = {k: 1e-3 for k in ("pos", "logscale", "quat", "logit", "sh")}
    # two views in one step (the same camera twice) sum their gradients
    p, m1, m2, hist = orc.train(g, [cam, cam], [view, view], [target, target], [[0, 1]], lrs)
    r = orc.render_and_grad(g, cam, view, target=target)
    for k in lrs:
        # the second view's terms are added one by one into the first view's
        # sums, so the total equals twice one view's up to float reordering
        assert np.abs(hist[0][1][k] - 2.0 * r["grads"][k]).max() <= 1e-5 * np.abs(r["grads"][k]).max(), k
        gk = r["grads"][k].astype(np.float64)
        mask = np.abs(gk) > 1e-6 * np.abs(gk).max()
        d = p[k].astype(np.float64) - g[k]
        # first step from zero moments: -lr g / (|g| + eps) = -lr sign(g)
        assert np.abs(d[mask] + 1e-3 * np.sign(gk[mask])).max() <= 1e-6, k
        assert np.allclose(m1[k], 0.1 * hist[0][1][k], rtol=1e-6, atol=1e-12)
