"""The 300-step training run on the device: 500,000 Gaussians, 8 ring views at 1280x720,
300 plain-method Adam iterations cycling the views, prune at
sigmoid(logit) < 0.005 and densify (split if the projected radius exceeds
8 px and a tile's L1 error is above its view's 90th percentile, else clone)
every 100 iterations.

At iterations 1, 101 and 201 the current view's sort keys, values, tile
offsets and image are compared with the oracle on the device's current
parameters; at every prune / densify event the survivor set, the decisions
and the compacted / grown arrays are compared with the oracle on the same
state. Targets are rendered by the device's forward (bit-identical to the
oracle's, see test_gpu_parity) from a jittered copy.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")
N, V, W, H, ITERS, EVERY = 500_000, 8, 1280, 720, 300, 100


def _check_view(c, orc, cam, v, tag):
    g = c.get_gaussians()
    view = orc.view_params(W, H, sh_deg=3)
    ref = orc.render_and_grad(g, cam, view)
    n = c.build_tiles(v)
    assert n == ref["n_isect"], tag
    keys, vals = c.keys(v, n)
    assert np.array_equal(keys, ref["keys"]) and np.array_equal(vals, ref["vals"]), tag
    assert np.array_equal(c.tile_offsets(v), ref["offsets"]), tag
    c.forward(v)
    err = float(np.abs(c.image(v) - ref["color"]).max())
    print(f"[config1] {tag}: I={n} keys/offsets exact, image max err {err:.2e}", flush=True)
    assert err <= 1e-5, tag


def test_config1_training_run(sv, orc):
    g = sv.scene_gaussians(0, N, 1.0)
    cams = sv.ring_cameras(V, 3.0, W, H)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    c = sv.Context(capacity=N)
    c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                 lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                 pos_lr_schedule=0)
    c.set_gaussians(tg)
    c.set_views(cams)
    targets = []
    for v in range(V):
        c.forward(v)
        targets.append(c.image(v))
    c.set_gaussians(g)
    for v in range(V):
        c.set_target(v, targets[v])
    t_steps = 0.0
    for it in range(1, ITERS + 1):
        v = (it - 1) % V
        if it in (1, 101, 201):
            _check_view(c, orc, cams[v], v, f"iteration {it} view {v}")
        t0 = time.perf_counter()
        c.step_views([v], sv.MODE_PLAIN)
        c.sync()
        t_steps += time.perf_counter() - t0
        if it % EVERY:
            continue
        # prune on the device vs the oracle's decision and compaction
        st, m1, m2 = c.get_gaussians(), c.get_moments(1), c.get_moments(2)
        keep = orc.prune_keep(st["logit"], 0.005).astype(bool)
        removed = c.prune_opacity(0.005)
        assert removed == int((~keep).sum())
        got, g1, g2 = c.get_gaussians(), c.get_moments(1), c.get_moments(2)
        for k in KEYS:
            assert np.array_equal(got[k], st[k][keep]), k
            assert np.array_equal(g1[k], m1[k][keep]) and np.array_equal(g2[k], m2[k][keep]), k
        # densify rule on the device vs the oracle on the pruned state
        act, nc, ns = c.densify_decide(list(range(V)), 0.9, 8.0)
        views = [orc.view_params(W, H, sh_deg=3) for _ in range(V)]
        projs, l1s = [], []
        for k in range(V):
            r = orc.render_and_grad(got, cams[k], views[k])
            projs.append(r["proj"])
            l1s.append(orc.tile_l1(views[k], r["color"], targets[k]))
            assert np.array_equal(c.tile_l1(k), l1s[-1]), k
        thr = np.array([orc.percentile_lower(x, 0.9) for x in l1s], np.float32)
        want = orc.densify_decide(projs, views, l1s, thr, 8.0)
        assert np.array_equal(act, want)
        n_before = c.count
        added = c.densify(act, it)
        ref, r1, r2 = orc.densify(got, g1, g2, act, it)
        assert c.count == ref["pos"].shape[0] == n_before + added
        now, n1, n2 = c.get_gaussians(), c.get_moments(1), c.get_moments(2)
        nk = int((act != 2).sum())
        for k in KEYS:
            assert np.array_equal(now[k][:nk], ref[k][:nk]), k
            assert np.array_equal(n1[k], r1[k]) and np.array_equal(n2[k], r2[k]), k
            d = np.abs(now[k].view(np.int32).astype(np.int64) - ref[k].view(np.int32).astype(np.int64))
            assert d.max() <= 1, k
        print(f"[config1] iteration {it}: pruned {removed}, cloned {nc}, split {ns}, N {n_before} -> {c.count}",
              flush=True)
    l1 = 0.0
    for v in range(V):
        c.forward(v)
        l1 += float(np.abs(c.image(v) - targets[v]).mean())
    print(f"[config1] {ITERS} plain steps in {t_steps:.2f} s ({ITERS / t_steps:.1f} steps/s incl. per-step sync), "
          f"final mean L1 {l1 / V:.5f}, N {c.count}", flush=True)
    c.close()
