"""The 64-view workload of uneven sizes (seeded sizes, width 320-1920 and
height 240-1080, cameras on a sphere of radius 4, means in [-2,2]^3), in two
forms:

1. at 1/6 of the resolution (widths 53-320, heights 40-180) with 100,000
   splats, where the oracle covers every view: the graph-captured batched
   step renders all 64 views within 1e-5 of the oracle (exactly, in
   practice); each view's own tile build gives the oracle's sort keys,
   values and tile offsets exactly; and the batched step equals 64
   sequential per-view passes bit for bit (parameters, moments, images);
2. at full size with 1,000,000 splats (about 188k tiles in one key space),
   through size-independent properties: the batched graph step equals the
   sequential step bit for bit, every batched image equals that view's
   single-view render, and the two smallest views match the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")


def _cfg(c, graph):
    c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                 lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                 pos_lr_schedule=0, use_graph=graph)


def _step(sv, g, cams, targets, mode, graph):
    V = len(cams)
    with sv.Context(capacity=g["pos"].shape[0]) as c:
        _cfg(c, graph)
        c.set_gaussians(g)
        c.set_views(cams)
        for v, t in enumerate(targets):
            c.set_target(v, t)
        st = c.step_views(list(range(V)), mode, stats=True)
        return ([c.image(v) for v in range(V)], c.get_gaussians(), c.get_moments(1), c.get_moments(2),
                (int(st.n_isect), int(st.key_bits), int(st.binned)))


def test_uneven64_reduced_vs_oracle(sv, orc):
    n = 100_000
    g = sv.scene_gaussians(0, n, 2.0)
    cams, sizes = sv.sphere_cameras(2, 64, 4.0, wmin=320 // 6, wmax=1920 // 6, hmin=240 // 6, hmax=1080 // 6)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    views = [orc.view_params(w, h, sh_deg=3) for w, h in sizes]
    targets = [orc.render_and_grad(tg, cam, v)["color"] for cam, v in zip(cams, views)]
    refs = [orc.render_and_grad(g, cam, v) for cam, v in zip(cams, views)]
    # per-view tile builds: keys, values, offsets exact
    with sv.Context(capacity=n) as c:
        _cfg(c, 0)
        c.set_gaussians(g)
        c.set_views(cams)
        for v, r in enumerate(refs):
            m = c.build_tiles(v)
            assert m == r["n_isect"], v
            keys, vals = c.keys(v, m)
            assert np.array_equal(keys, r["keys"]) and np.array_equal(vals, r["vals"]), v
            assert np.array_equal(c.tile_offsets(v), r["offsets"]), v
    bq = _step(sv, g, cams, targets, sv.MODE_SEQUENTIAL, 0)
    bb = _step(sv, g, cams, targets, sv.MODE_BATCHED, 1)
    print(f"[uneven64/6] {sum(w * h for w, h in sizes)} pixels, batched I={bb[4][0]} key bits {bb[4][1]} "
          f"binned {bb[4][2]}", flush=True)
    assert bq[4][0] == bb[4][0] == sum(r["n_isect"] for r in refs)
    worst = 0.0
    for v, r in enumerate(refs):
        worst = max(worst, float(np.abs(bb[0][v] - r["color"]).max()))
        assert np.array_equal(bq[0][v], bb[0][v]), v
    print(f"[uneven64/6] worst image error vs oracle {worst:.2e}", flush=True)
    assert worst <= 1e-5
    for k in KEYS:
        for a, b in zip(bq[1:4], bb[1:4]):
            assert np.array_equal(a[k], b[k]), k


def test_uneven64_fullsize_properties(sv, orc):
    n = 1_000_000
    g = sv.scene_gaussians(0, n, 2.0)
    cams, sizes = sv.sphere_cameras(2, 64, 4.0)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    targets, singles = [], []
    with sv.Context(capacity=n) as c:
        _cfg(c, 0)
        c.set_gaussians(tg)
        c.set_views(cams)
        for v in range(64):
            c.forward(v)
            targets.append(c.image(v))
        c.set_gaussians(g)
        for v in range(64):
            c.forward(v)
            singles.append(c.image(v))
    bb = _step(sv, g, cams, targets, sv.MODE_BATCHED, 1)
    bq = _step(sv, g, cams, targets, sv.MODE_SEQUENTIAL, 0)
    tiles = sum(((w + 15) // 16) * ((h + 15) // 16) for w, h in sizes)
    print(f"[uneven64] {tiles} tiles, I={bb[4][0]}, key bits {bb[4][1]}, binned {bb[4][2]}", flush=True)
    assert bb[4][0] == bq[4][0]
    for v in range(64):
        assert np.array_equal(bb[0][v], singles[v]), v
        assert np.array_equal(bq[0][v], singles[v]), v
    for k in KEYS:
        for a, b in zip(bq[1:4], bb[1:4]):
            assert np.array_equal(a[k], b[k]), k
    for v in sorted(range(64), key=lambda i: sizes[i][0] * sizes[i][1])[:2]:
        w, h = sizes[v]
        ref = orc.render_and_grad(g, cams[v], orc.view_params(w, h, sh_deg=3))["color"]
        err = float(np.abs(singles[v] - ref).max())
        print(f"[uneven64] view {v} ({w}x{h}) vs oracle {err:.2e}", flush=True)
        assert err <= 1e-5
