"""The device path against the committed golden fixtures of the adjoint-check
scene (4,096 Gaussians, 64x64, one view; tests/golden): sort keys, values,
tile offsets, alpha and last indices bit-exact, image within 1e-5, the
gradients of the 64 recorded Gaussians within 1e-4 of max |g|."""
import os
import sys

import numpy as np
import pytest

from conftest import close_rel

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as mg  # noqa: E402


def test_config2_against_golden(sv):
    z = np.load(os.path.join(HERE, "golden", "adjoint64_4096_64x64.npz"))
    n, w, h, eye = mg.SCENES["adjoint64"]
    g = sv.scene_gaussians(0, n, 1.0)
    cam = sv.scene_camera(list(eye), [0, 0, 0], w, h, 60.0)
    c = sv.Context(capacity=n)
    c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0)
    c.set_gaussians(g)
    c.set_views([cam])
    c.set_target(0, z["target"])
    m = c.build_tiles(0)
    keys, vals = c.keys(0, m)
    assert np.array_equal(keys, z["keys"]) and np.array_equal(vals, z["vals"])
    assert np.array_equal(c.tile_offsets(0), z["offsets"])
    c.forward(0)
    assert np.abs(c.image(0) - z["color"]).max() <= 1e-5
    assert np.array_equal(c.alpha(0).view(np.uint32), z["alpha"].view(np.uint32))
    assert np.array_equal(c.last_ids(0), z["last_ids"])
    c.zero_grad()
    c.backward(0)
    gg = c.grads()
    pick = z["pick"]
    for k in mg.GRADS:
        e = close_rel(gg[k][pick], z["grad_" + k], 1e-4)
        print(f"[golden] adjoint64 grad {k}: {e:.2e}", flush=True)
        assert e <= 1e-4, k
    c.close()
