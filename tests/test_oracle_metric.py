"""The oracle's multi-view scoring (orc_metric_view + multiview_scores:
metric-map.wgsl:32-147, gaussian-renderer.ts:2709-2781) against numpy: the
l1 metric map, min / max / mean, the threshold flags, and the per-splat
counts by a brute-force walk of the flagged pixels' tile lists; invariants:
a splat that colours a flagged pixel (alpha >= 1/255, so power <= ln 255 <
6) is counted there."""
import numpy as np


def _scene(sv, orc, n=2000, w=64, h=48, seed=4):
    g = sv.scene_gaussians(seed, n, 1.0)
    cam = sv.scene_camera([0.3, 0.2, 3.0], [0, 0, 0], w, h, 60.0)
    view = orc.view_params(w, h, sh_deg=3)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(seed + 1, g["pos"], 0.05)
    target = orc.render_and_grad(tg, cam, view)["color"]
    return g, cam, view, target


def _brute(r, view, target, n):
    col = r["color"].astype(np.float32)
    l1 = ((np.abs(col[..., 0] - target[..., 0]) + np.abs(col[..., 1] - target[..., 1]))
          + np.abs(col[..., 2] - target[..., 2])) / np.float32(3)
    l1 = l1.astype(np.float32)
    mn, mx = float(l1.min()), float(l1.max())
    inv = np.float32(1.0 / (mx - mn)) if mx > mn else np.float32(0)
    flags = ((l1 - np.float32(mn)) * inv) > np.float32(0.1)
    counts = np.zeros(n, np.int64)
    proj = r["proj"]
    H, W = l1.shape
    for y, x in zip(*np.nonzero(flags)):
        t = (y // 16) * view.tiles_x + x // 16
        s, e = int(r["offsets"][t]), int(r["offsets"][t + 1])
        last = int(r["last_ids"][y, x])
        if s >= e or last < s:
            continue
        for i in range(s, min(last, e - 1) + 1):
            gi = int(r["vals"][i])
            m = proj["mean2d"][gi]
            a, b, c = proj["conic"][gi]
            dx = np.float32(x + 0.5) - m[0]
            dy = np.float32(y + 0.5) - m[1]
            p = np.float32(0.5) * (a * dx * dx + c * dy * dy) + b * dx * dy
            if 0 <= p <= 6:
                counts[gi] += 1
    return l1, flags, counts


def test_metric_view_against_brute_force(sv, orc):
    g, cam, view, target = _scene(sv, orc)
    n = g["pos"].shape[0]
    imp, pru, scored = orc.multiview_scores(g, [cam], [view], [target])
    r = orc.render_and_grad(g, cam, view)
    l1, flags, counts = _brute(r, view, target, n)
    assert scored == 1 and 0 < flags.sum() < flags.size
    assert np.array_equal(imp.astype(np.int64), counts)
    mean = float(l1.astype(np.float64).mean())
    assert np.allclose(pru, (counts * mean).astype(np.float32), rtol=1e-6, atol=0)


def test_metric_counts_cover_contributors(sv, orc):
    """With every pixel flagged (a target that differs everywhere but at one
    pixel), each splat is counted at least at every pixel it colours."""
    g, cam, view, _ = _scene(sv, orc, seed=9)
    n = g["pos"].shape[0]
    r = orc.render_and_grad(g, cam, view)
    target = r["color"] + np.float32(0.5)
    target[0, 0] = r["color"][0, 0]  # the minimum: every other pixel flagged
    imp, _, _ = orc.multiview_scores(g, [cam], [view], [target])
    proj = r["proj"]
    contrib = np.zeros(n, np.int64)
    H, W = r["alpha"].shape
    for y in range(H):
        for x in range(W):
            if x == 0 and y == 0:
                continue
            t = (y // 16) * view.tiles_x + x // 16
            s, e = int(r["offsets"][t]), int(r["offsets"][t + 1])
            if s >= e:
                continue
            for i in range(s, min(int(r["last_ids"][y, x]), e - 1) + 1):
                gi = int(r["vals"][i])
                m = proj["mean2d"][gi]
                a, b, c = proj["conic"][gi]
                dx, dy = np.float32(x + 0.5) - m[0], np.float32(y + 0.5) - m[1]
                p = np.float32(0.5) * (a * dx * dx + c * dy * dy) + b * dx * dy
                op = 1.0 / (1.0 + np.exp(-float(g["logit"][gi])))
                if p >= 0 and op * np.exp(-float(p)) >= 1.0 / 255 + 1e-6:
                    contrib[gi] += 1
    assert contrib.sum() > 0
    assert np.all(imp.astype(np.int64) >= contrib)
