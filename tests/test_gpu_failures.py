"""Failure paths forced from the host (each returns a status code; every
kernel still runs to its normal end):

- no Gaussians: steps, renders and tile builds return SV_E_STATE, prune and
  densify do nothing;
- NaN / infinite inputs: parameters, moments, targets and residuals with a
  non-finite value return SV_E_ARG and leave the state as it was;
- more Gaussians than the capacity: SV_E_CAPACITY;
- fixed-point overflow in the reference's method: with a loss weight of
  3e4 the per-intersection i32 sums of backward-tiles-local.wgsl pass
  2^31 and wrap (two's complement, as the shader's atomicAdd on i32 does,
  behaviour A13); the device's integers equal the oracle's bit for bit, and
  differ from a linear rescale of the unscaled loss's, so wrapping occurs.
"""
import numpy as np
import pytest

from conftest import small_scene

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")


def _status(fn, *a):
    import sv

    with pytest.raises(sv.SvError) as e:
        fn(*a)
    return e.value.status


def test_zero_gaussians(sv):
    g = sv.scene_gaussians(0, 10, 1.0)
    empty = {k: v[:0] for k, v in g.items()}
    cam = sv.scene_camera([0.0, 0.0, 3.0], [0, 0, 0], 64, 64, 60.0)
    with sv.Context(capacity=16) as c:
        c.set_gaussians(empty)
        c.set_views([cam])
        c.set_target(0, np.zeros((64, 64, 3), np.float32))
        assert c.count == 0
        assert _status(c.step_views, [0], sv.MODE_BATCHED) == sv.SV_E_STATE
        assert _status(c.step_views, [0], sv.MODE_PLAIN) == sv.SV_E_STATE
        assert _status(c.render_camera, cam) == sv.SV_E_STATE
        assert _status(c.render_camera_noalloc, cam) == sv.SV_E_STATE
        assert _status(c.build_tiles, 0) == sv.SV_E_STATE
        assert c.prune_opacity(0.5) == 0
        assert c.densify(np.zeros(0, np.uint8), 1) == 0
        # the context still works afterwards
        c.set_gaussians(g)
        c.step_views([0], sv.MODE_BATCHED)
        assert np.isfinite(c.get_gaussians()["pos"]).all()


def test_nonfinite_inputs_rejected(sv):
    g, tg, cam = small_scene(sv, n=4096, w=64, h=64)
    with sv.Context(capacity=4096) as c:
        c.set_gaussians(g)
        c.set_views([cam])
        good_t = np.full((64, 64, 3), 0.25, np.float32)
        c.set_target(0, good_t)
        c.step_views([0], sv.MODE_BATCHED)
        before, m1 = c.get_gaussians(), c.get_moments(1)
        for key, val in (("logit", np.nan), ("pos", np.inf), ("sh", -np.inf), ("quat", np.nan), ("logscale", np.nan)):
            bad = {k: v.copy() for k, v in g.items()}
            bad[key].reshape(-1)[17] = val
            assert _status(c.set_gaussians, bad) == sv.SV_E_ARG, key
            assert "not finite" in sv.lib().sv_last_error(c.h).decode()
            badm = {k: v.copy() for k, v in m1.items()}
            badm[key].reshape(-1)[3] = val
            assert _status(c.set_moments, 1, badm) == sv.SV_E_ARG, key
        bt = good_t.copy()
        bt[5, 7, 1] = np.nan
        assert _status(c.set_target, 0, bt) == sv.SV_E_ARG
        assert _status(c.set_residual, 0, bt) == sv.SV_E_ARG
        after, m1b = c.get_gaussians(), c.get_moments(1)
        for k in KEYS:
            assert np.array_equal(before[k], after[k]) and np.array_equal(m1[k], m1b[k]), k
        # training goes on from the unchanged state
        c.step_views([0], sv.MODE_BATCHED)
        assert np.isfinite(c.get_gaussians()["pos"]).all()


def test_capacity_exceeded(sv):
    g = sv.scene_gaussians(0, 100, 1.0)
    with sv.Context(capacity=64) as c:
        assert _status(c.set_gaussians, g) == sv.SV_E_CAPACITY
        assert c.count == 0


@pytest.mark.parametrize("flags", [0, 3])
def test_fixed_point_overflow_wraps_as_the_oracle(sv, orc, flags):
    g, tg, cam = small_scene(sv)
    view = orc.view_params(256, 256, sh_deg=3)
    target = orc.render_and_grad(tg, cam, view)["color"]
    big = orc.render_and_grad(g, cam, view, target=target, w_l2=3e4, w_l1=0.0, wavefront=True, wf_flags=flags)
    unit = orc.render_and_grad(g, cam, view, target=target, w_l2=1.0, w_l1=0.0, wavefront=True, wf_flags=flags)
    fb = big["tile_grads_fixed"].astype(np.int64)
    fu = unit["tile_grads_fixed"].astype(np.int64)
    # without wrapping the sums would scale with the weight (up to the 2^-16
    # truncation of each term); with it many are far off
    off = np.abs(fb - 30000 * fu) > 30000 * 256 + 2**20
    print(f"[overflow] flags {flags}: {int(off.sum())} of {off.size} fixed-point sums wrapped, "
          f"max |int| {int(np.abs(fb).max())}", flush=True)
    assert off.sum() > 0
    with sv.Context(capacity=32768) as c:
        c.set_config(sh_deg=3, w_l2=3e4, w_l1=0.0, w_ssim=0.0)
        c.set_gaussians(g)
        c.set_views([cam])
        c.set_target(0, target)
        c.forward(0)
        c.zero_grad()
        c.backward_fixed(0, flags)
        got = c.tile_grads_fixed(0, big["n_isect"])
    assert np.array_equal(got, big["tile_grads_fixed"])
