"""A batched growth run past the context's initial capacity (the shape of the
1M -> 4M growth workload at a size the oracle checks in seconds): 100,000
Gaussians in a context of exactly that capacity, 4 ring views at 480x270,
batched adjoint steps launched as a captured graph, and three prune /
densify events (every 10 steps) with the reference's growth cap of
max(1, floor(0.1 N)) new splats per event (gaussian-renderer.ts:3297-3306).

At every event, on the device's state: the prune survivors, the capped
densify decisions (scores and cap order included) and the grown arrays
equal the oracle's; the scene outgrows the initial capacity at the first
event. After every event the next batched step (graph re-captured on the
grown buffers) gives the same parameters, bit for bit, as a sequential
step from the same state, and its images equal the oracle's render of that
state within 1e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")
N0, V, W, H, EVERY, EVENTS = 100_000, 4, 480, 270, 10, 3


def _cfg(c, graph):
    c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                 lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                 pos_lr_schedule=0, use_graph=graph)


def test_batched_growth_past_capacity(sv, orc):
    g = sv.scene_gaussians(0, N0, 1.0)
    g["logscale"] = (g["logscale"] + np.float32(0.7)).astype(np.float32)
    cams = sv.ring_cameras(V, 3.0, W, H)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    views = [orc.view_params(W, H, sh_deg=3) for _ in range(V)]
    targets = [orc.render_and_grad(tg, cam, v)["color"] for cam, v in zip(cams, views)]
    c = sv.Context(capacity=N0)
    _cfg(c, 1)
    c.set_gaussians(g)
    c.set_views(cams)
    for v, t in enumerate(targets):
        c.set_target(v, t)
    vs = list(range(V))
    grew = False
    try:
        for ev in range(1, EVENTS + 1):
            for _ in range(EVERY):
                c.step_views(vs, sv.MODE_BATCHED)
            # prune
            st, m1, m2 = c.get_gaussians(), c.get_moments(1), c.get_moments(2)
            keep = orc.prune_keep(st["logit"], 0.005).astype(bool)
            removed = c.prune_opacity(0.005)
            assert removed == int((~keep).sum())
            got, g1, g2 = c.get_gaussians(), c.get_moments(1), c.get_moments(2)
            for k in KEYS:
                assert np.array_equal(got[k], st[k][keep]), k
            # capped densify decision on the pruned state
            n_before = c.count
            max_new = max(1, n_before // 10)
            act, nc, ns = c.densify_decide(vs, 0.9, 8.0, max_new=max_new)
            projs, l1s = [], []
            for k in range(V):
                r = orc.render_and_grad(got, cams[k], views[k])
                projs.append(r["proj"])
                l1s.append(orc.tile_l1(views[k], r["color"], targets[k]))
            thr = np.array([orc.percentile_lower(x, 0.9) for x in l1s], np.float32)
            want, score = orc.densify_decide(projs, views, l1s, thr, 8.0, with_score=True)
            want = orc.densify_cap(want, score, max_new)
            assert np.array_equal(act, want)
            added = c.densify(act, ev)
            ref, r1, r2 = orc.densify(got, g1, g2, act, ev)
            assert c.count == ref["pos"].shape[0] == n_before + added
            assert added <= max_new
            now, n1, n2 = c.get_gaussians(), c.get_moments(1), c.get_moments(2)
            nk = int((act != 2).sum())
            for k in KEYS:
                assert np.array_equal(now[k][:nk], ref[k][:nk]), k
                assert np.array_equal(n1[k], r1[k]) and np.array_equal(n2[k], r2[k]), k
                d = np.abs(now[k].view(np.int32).astype(np.int64) - ref[k].view(np.int32).astype(np.int64))
                assert d.max() <= 1, k
            grew = grew or c.count > N0
            print(f"[growth] event {ev}: pruned {removed}, cloned {nc}, split {ns} (cap {max_new}), "
                  f"N {n_before} -> {c.count}", flush=True)
            # the next batched (graph) step = a sequential step from the same state
            state = (c.get_gaussians(), c.get_moments(1), c.get_moments(2), c.step)
            with sv.Context(capacity=c.count) as q:
                _cfg(q, 0)
                q.set_gaussians(state[0])
                q.set_moments(1, state[1])
                q.set_moments(2, state[2])
                q.set_views(cams)
                for v, t in enumerate(targets):
                    q.set_target(v, t)
                lib = sv.lib()
                assert lib.sv_set_step(q.h, state[3]) == 0
                q.step_views(vs, sv.MODE_SEQUENTIAL)
                seq = q.get_gaussians()
            c.step_views(vs, sv.MODE_BATCHED)
            after = c.get_gaussians()
            for k in KEYS:
                assert np.array_equal(after[k], seq[k]), (ev, k)
            err = max(float(np.abs(c.image(v) - orc.render_and_grad(state[0], cams[v], views[v])["color"]).max())
                      for v in range(V))
            assert err <= 1e-5, err
        assert grew, "the run must outgrow the initial capacity"
    finally:
        c.close()
