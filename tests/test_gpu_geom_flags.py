"""SV_GEOM_MEAN_Y_FORWARD: the geometry backward's mean path takes the
forward's y convention (pixel y = H - c_y - f_y y / z) instead of the
shader's +f_y (training-backward-geom.wgsl:264-266, whose position-y
gradient has the opposite sign to the forward's dependence). On the device
both geometry kernels (k_vjp_geom of the adjoint modes, k_geom of the plain
mode) match the oracle run with the same flag within 1e-4 of max; the flag
changes the position gradients only (logit, SH, 2-D geometry, log-scale and
quaternion gradients are bit-identical with and without it).
"""
import numpy as np
import pytest

from conftest import small_scene, close_rel

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-4


def _ctx(sv, g, cam, target, flags, **kw):
    c = sv.Context(capacity=32768)
    c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                 lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                 pos_lr_schedule=0, geom_flags=flags, **kw)
    c.set_gaussians(g)
    c.set_views([cam])
    c.set_target(0, target)
    return c


@pytest.fixture(scope="module")
def scene(sv, orc):
    g, tg, cam = small_scene(sv)
    target = orc.render_and_grad(tg, cam, orc.view_params(256, 256, sh_deg=3))["color"]
    return g, cam, target


def test_adjoint_geometry_with_forward_y(sv, orc, scene):
    g, cam, target = scene
    ref = orc.render_and_grad(g, cam, orc.view_params(256, 256, sh_deg=3, geom_flags=1), target=target,
                              w_l2=0.8, w_l1=0.2)
    got = {}
    for flags in (0, sv.GEOM_MEAN_Y_FORWARD):
        with _ctx(sv, g, cam, target, flags) as c:
            c.forward(0)
            c.zero_grad()
            c.backward(0)
            got[flags] = c.grads()
    gg = got[sv.GEOM_MEAN_Y_FORWARD]
    for k in ("pos", "logscale", "quat", "logit", "sh", "geom"):
        e = close_rel(gg[k], ref["grads"][k], GRAD_TOL)
        print(f"[geom-y] adjoint grad {k}: {e:.3e}", flush=True)
        assert e <= GRAD_TOL, k
    for k in ("logscale", "quat", "logit", "sh", "geom"):
        assert np.array_equal(got[0][k], gg[k]), k
    assert not np.array_equal(got[0]["pos"], gg["pos"])


def test_plain_geometry_with_forward_y(sv, orc, scene):
    g, cam, target = scene
    ref = orc.render_and_grad(g, cam, orc.view_params(256, 256, sh_deg=3, geom_flags=1), target=target,
                              w_l2=0.8, w_l1=0.2, wavefront=True, wf_flags=0)
    with _ctx(sv, g, cam, target, sv.GEOM_MEAN_Y_FORWARD, plain_wf_flags=0) as c:
        c.step_views([0], sv.MODE_PLAIN)
        gg = c.grads()
    for k in ("pos", "logscale", "quat", "logit", "sh"):
        e = close_rel(gg[k], ref["grads"][k], GRAD_TOL)
        print(f"[geom-y] plain grad {k}: {e:.3e}", flush=True)
        assert e <= GRAD_TOL, k


def test_unknown_geometry_flag_rejected(sv, scene):
    g, cam, target = scene
    with sv.Context(capacity=1024) as c:
        with pytest.raises(sv.SvError) as e:
            c.set_config(geom_flags=2)
        assert e.value.status == sv.SV_E_ARG
