"""render_f16 (the reference's f16 render copy, training-params.wgsl:81-128,
behaviour A6): with the option on, every pass that reads parameters
(projection, SH colour, the fixed-point backward's SH, the geometry and SH
VJPs) reads them rounded to f16 (RNE), while Adam updates the f32 masters.

Parameters that are not f16-representable (the seeded scene plus N(0, 1e-3)
noise) are trained for one step on the device and compared with the oracle
run on the f16-rounded parameters (numpy float16 RNE, which also pins the
oracle's rounding): the image bit for bit, gradients within 1e-4 of max |ref|
in the adjoint and the reference-method (plain) modes, and the new
parameters equal to Adam applied to the unrounded masters with the device's
own gradients. The unrounded scene renders a different image, so the
rounding is really exercised.
"""
import numpy as np
import pytest

from conftest import close_rel, small_scene

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")


def _cfg(ctx, **kw):
    ctx.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                   lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                   pos_lr_schedule=0, render_f16=1, **kw)


@pytest.mark.parametrize("mode", ["sequential", "plain"])
def test_render_f16_step(sv, orc, mode):
    g, tg, cam = small_scene(sv)
    rng = np.random.default_rng(11)
    raw = {k: (g[k] + rng.normal(0.0, 1e-3, g[k].shape)).astype(np.float32) for k in KEYS}
    rounded = {k: raw[k].astype(np.float16).astype(np.float32) for k in KEYS}
    assert np.array_equal(orc.f16_round(raw["pos"]), rounded["pos"])
    assert not np.array_equal(raw["pos"], rounded["pos"])
    view = orc.view_params(256, 256, sh_deg=3)
    target = orc.render_and_grad(tg, cam, view)["color"]
    wf = mode == "plain"
    ref = orc.render_and_grad(rounded, cam, view, target=target, w_l2=0.8, w_l1=0.2, wavefront=wf, wf_flags=0)
    unrounded = orc.render_and_grad(raw, cam, view)["color"]
    assert not np.array_equal(unrounded, ref["color"])
    with sv.Context(capacity=32768) as c:
        _cfg(c, plain_wf_flags=0)
        c.set_gaussians(raw)
        c.set_views([cam])
        c.set_target(0, target)
        c.step_views([0], sv.MODE_PLAIN if wf else sv.MODE_SEQUENTIAL)
        img = c.image(0)
        gr = c.grads()
        got = c.get_gaussians()
    err = float(np.abs(img - ref["color"]).max())
    print(f"[f16] {mode}: image err {err:.3e}", flush=True)
    assert np.array_equal(img, ref["color"])
    bc1, bc2 = orc.adam_bias(1, 0.9, 0.999)
    for k in KEYS:
        e = close_rel(gr[k], ref["grads"][k], 1e-4)
        print(f"[f16] {mode}: grad {k} {e:.3e}", flush=True)
        assert e <= 1e-4, k
        p = raw[k].copy()
        orc.adam(p, np.zeros_like(p), np.zeros_like(p), gr[k], 1e-3, 0.9, 0.999, 1e-15, bc1, bc2)
        assert np.array_equal(got[k], p), k
