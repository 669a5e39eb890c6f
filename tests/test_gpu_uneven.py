"""Views of uneven sizes in one batched pipeline (the 64-view workload at reduced size:
seeded sizes, cameras on a sphere of radius 4, means in [-2,2]^3): every
image matches the oracle within 1e-5 (exactly, in practice), and the
batched step — captured as a graph — is bit-identical to per-view
sequential adjoint passes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")


def test_uneven_views_batched(sv, orc):
    n = 30000
    g = sv.scene_gaussians(0, n, 2.0)
    cams, sizes = sv.sphere_cameras(2, 6, 4.0, wmin=96, wmax=400, hmin=64, hmax=300)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    views = [orc.view_params(w, h, sh_deg=3) for w, h in sizes]
    targets = [orc.render_and_grad(tg, cam, v)["color"] for cam, v in zip(cams, views)]
    out = {}
    for mode, graph in ((sv.MODE_SEQUENTIAL, 0), (sv.MODE_BATCHED, 1)):
        c = sv.Context(capacity=n)
        c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                     lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                     pos_lr_schedule=0, keep_grads=1, use_graph=graph)
        c.set_gaussians(g)
        c.set_views(cams)
        for k, t in enumerate(targets):
            c.set_target(k, t)
        st = c.step_views(list(range(len(cams))), mode, stats=True)
        out[mode] = ([c.image(k) for k in range(len(cams))], c.grads(), c.get_gaussians(), st)
        c.close()
    (iq, gq, pq, sq), (ib, gb, pb, sb) = out[sv.MODE_SEQUENTIAL], out[sv.MODE_BATCHED]
    print(f"[uneven] sizes {sizes}; batched I={sb.n_isect}, key bits {sb.key_bits}, passes {sb.sort_passes}",
          flush=True)
    assert sq.n_isect == sb.n_isect
    for k, (cam, v) in enumerate(zip(cams, views)):
        ref = orc.render_and_grad(g, cam, v)["color"]
        assert np.abs(ib[k] - ref).max() <= 1e-5, k
        assert np.array_equal(iq[k], ib[k]), k
    for k in KEYS:
        assert np.array_equal(gq[k], gb[k]) and np.array_equal(pq[k], pb[k]), k
