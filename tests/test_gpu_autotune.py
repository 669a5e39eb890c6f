"""The autotuner (sv_autotune) times every candidate, leaves the fastest
configured, and leaves the training state as it found it: a step after
tuning equals the same step on a fresh context with the chosen
configuration, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")


def _ctx(sv, g, cams, targets, **kw):
    c = sv.Context(capacity=32768)
    c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                 lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                 pos_lr_schedule=0, **kw)
    c.set_gaussians(g)
    c.set_views(cams)
    for k, t in enumerate(targets):
        c.set_target(k, t)
    return c


def test_autotune_restores_state_and_keeps_the_fastest(sv, orc):
    g = sv.scene_gaussians(0, 16384, 1.0)
    cams = sv.ring_cameras(4, 3.0, 256, 192)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    targets = [orc.render_and_grad(tg, cam, orc.view_params(256, 192, sh_deg=3))["color"] for cam in cams]
    c = _ctx(sv, g, cams, targets, use_graph=1)
    c.step_views([0, 1, 2, 3], sv.MODE_BATCHED)  # non-trivial moments and step counter
    before, m1 = c.get_gaussians(), c.get_moments(1)
    table = c.autotune([0, 1, 2, 3], 2)
    print("[tune] " + ", ".join(f"t{e['tile']}/b{e['batch']}/p{e['persistent']}/x{e['xcd_remap']} "
                                f"{e['ms_per_step']:.3f} ms" for e in table), flush=True)
    # raster tile 16: batch x persistent x XCD mapping (12); 8 and 32: batch x XCD mapping (6 each); then
    # the sweep's fastest with the split backward switched (four views)
    assert len(table) == 25 and all(e["ms_per_step"] > 0 for e in table)
    assert sorted({e["tile"] for e in table}) == [8, 16, 32]
    split0 = table[0]["split_backward"]
    assert all(e["split_backward"] == split0 for e in table[:24])
    sweep_best = min(table[:24], key=lambda e: e["ms_per_step"])
    last = table[24]
    assert last["split_backward"] == 1 - split0
    assert (last["tile"], last["batch"], last["persistent"], last["xcd_remap"]) == (
        sweep_best["tile"], sweep_best["batch"], sweep_best["persistent"], sweep_best["xcd_remap"])
    best = min(table, key=lambda e: e["ms_per_step"])
    cfg = c.config()
    assert (cfg.raster_px, cfg.batch, cfg.persistent, cfg.xcd_remap, cfg.split_backward) == (
        best["tile"], best["batch"], best["persistent"], best["xcd_remap"], best["split_backward"])
    after, m1b = c.get_gaussians(), c.get_moments(1)
    for k in KEYS:
        assert np.array_equal(before[k], after[k]) and np.array_equal(m1[k], m1b[k]), k
    c.step_views([0, 1, 2, 3], sv.MODE_BATCHED)
    tuned = c.get_gaussians()
    c.close()
    f = _ctx(sv, g, cams, targets, use_graph=1, batch=best["batch"], persistent=best["persistent"],
             xcd_remap=best["xcd_remap"], raster_px=best["tile"], split_backward=best["split_backward"])
    f.step_views([0, 1, 2, 3], sv.MODE_BATCHED)
    f.step_views([0, 1, 2, 3], sv.MODE_BATCHED)
    fresh = f.get_gaussians()
    f.close()
    for k in KEYS:
        assert np.array_equal(tuned[k], fresh[k]), k
