"""Pins of the oracle's prune / densify / scoring restatements (CPU only).

The reference decides by weighted random scores and draws Math.random, so
nothing of it can be reproduced bit for bit; these functions follow its
compaction layout (gaussian-renderer.ts:2917-3000, 3211-3466) with the
seeded counter stream, and are pinned here by independent numpy
restatements of the layout and the draws, float64 sums and brute force.
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as mg  # noqa: E402


def _rand_params(rng, n):
    q = rng.normal(size=(n, 4)).astype(np.float32)
    return {"pos": rng.uniform(-1, 1, (n, 3)).astype(np.float32),
            "logscale": rng.normal(-4.5, 0.5, (n, 3)).astype(np.float32),
            "quat": q,
            "logit": rng.normal(0, 2, n).astype(np.float32),
            "sh": rng.normal(0, 0.1, (n, 48)).astype(np.float32)}


def test_prune_keep_matches_float64_sigmoid(orc):
    rng = np.random.default_rng(0)
    lg = rng.normal(-5.3, 1.0, 200000).astype(np.float32)
    keep = orc.prune_keep(lg, 0.005)
    sig = 1.0 / (1.0 + np.exp(-lg.astype(np.float64)))
    far = np.abs(sig - 0.005) > 1e-8
    assert np.array_equal(keep[far].astype(bool), sig[far] >= 0.005)
    assert 0 < keep.sum() < keep.size


def test_densify_layout_and_draws(orc):
    rng = np.random.default_rng(1)
    n = 2000
    g = _rand_params(rng, n)
    m1 = _rand_params(rng, n)
    m2 = _rand_params(rng, n)
    act = rng.choice([0, 0, 0, 1, 2], n).astype(np.uint8)
    seed = 987654321
    out, o1, o2 = orc.densify(g, m1, m2, act, seed)
    rows, r1, r2 = orc.to_rows(g), orc.to_rows(m1), orc.to_rows(m2)
    O, O1, O2 = orc.to_rows(out), orc.to_rows(o1), orc.to_rows(o2)
    surv = act != 2
    ns = int(surv.sum())
    assert O.shape[0] == ns + int((act == 1).sum()) + 2 * int((act == 2).sum())
    assert np.array_equal(O[:ns], rows[surv]) and np.array_equal(O1[:ns], r1[surv])
    assert np.array_equal(O2[:ns], r2[surv])
    assert not O1[ns:].any() and not O2[ns:].any()
    w = ns
    samples = []
    for i in np.flatnonzero(act):
        base = 16 * int(i)
        u = lambda k: float(mg.uniform(seed, np.uint64(base + k)))  # noqa: E731
        p = rows[i, 0:3].astype(np.float64)
        ls = rows[i, 3:6].astype(np.float64)
        for c in range(1 if act[i] == 1 else 2):
            child = O[w]
            assert np.array_equal(child[6:], rows[i, 6:])  # quaternion, logit, SH
            if act[i] == 1:
                mag = np.exp(ls).max() * 0.25 * (0.5 + u(0) * 0.5)
                d = np.array([u(1) * 2 - 1, u(2) * 2 - 1, u(3) * 2 - 1])
                want = (p + d / np.linalg.norm(d) * mag).astype(np.float32)
                assert np.array_equal(child[3:6], rows[i, 3:6])
                assert np.abs(child[0:3].astype(np.float64) - want).max() <= 1e-7
            else:
                r1_, r3_ = math.sqrt(-2 * math.log(1 - u(4 * c + 0))), math.sqrt(-2 * math.log(1 - u(4 * c + 2)))
                z = np.array([r1_ * math.cos(2 * math.pi * u(4 * c + 1)), r1_ * math.sin(2 * math.pi * u(4 * c + 1)),
                              r3_ * math.cos(2 * math.pi * u(4 * c + 3))])
                x, y, zq, wq = rows[i, 6:10].astype(np.float64)
                R = np.array([[1 - 2 * (y * y + zq * zq), 2 * (x * y - wq * zq), 2 * (x * zq + wq * y)],
                              [2 * (x * y + wq * zq), 1 - 2 * (x * x + zq * zq), 2 * (y * zq - wq * x)],
                              [2 * (x * zq - wq * y), 2 * (y * zq + wq * x), 1 - 2 * (x * x + y * y)]])
                want = (p + R @ (z * np.exp(ls))).astype(np.float32)
                assert np.abs(child[0:3].astype(np.float64) - want).max() <= 1e-6
                assert np.array_equal(child[3:6], (ls - math.log(1.6)).astype(np.float32))
                samples.append(z)
            w += 1
    assert w == O.shape[0]
    z = np.array(samples)
    assert np.abs(z.mean(0)).max() < 0.1 and np.abs(z.std(0) - 1).max() < 0.1


def test_tile_l1_tree_and_sum(orc):
    rng = np.random.default_rng(2)
    W, H = 70, 45  # ragged edge tiles
    view = orc.view_params(W, H)
    c = rng.uniform(0, 1, (H, W, 3)).astype(np.float32)
    t = rng.uniform(0, 1, (H, W, 3)).astype(np.float32)
    got = orc.tile_l1(view, c, t)
    e = np.abs(c - t)
    e = (e[..., 0] + e[..., 1]) + e[..., 2]
    for tid in range(view.tiles_x * view.tiles_y):
        tx, ty = tid % view.tiles_x, tid // view.tiles_x
        slot = np.zeros(256, np.float32)
        for k in range(256):
            x, y = tx * 16 + (k & 15), ty * 16 + (k >> 4)
            if x < W and y < H:
                slot[k] = e[y, x]
        exact = float(slot.astype(np.float64).sum())
        s = 128
        while s:
            slot[:s] = slot[:s] + slot[s:2 * s]
            s >>= 1
        assert got[tid] == slot[0]
        assert abs(got[tid] - exact) <= 1e-5 * max(exact, 1.0)


def test_densify_decide_brute_force(sv, orc):
    g = sv.scene_gaussians(3, 3000, 1.0)
    g["logscale"] = (g["logscale"] + np.float32(2.6)).astype(np.float32)
    cams = sv.ring_cameras(2, 3.0, 96, 64)
    views = [orc.view_params(96, 64, sh_deg=3) for _ in cams]
    projs = [orc.project(g, cam, v) for cam, v in zip(cams, views)]
    rng = np.random.default_rng(4)
    l1s = [rng.uniform(0, 10, v.tiles_x * v.tiles_y).astype(np.float32) for v in views]
    thr = np.array([orc.percentile_lower(x, 0.9) for x in l1s], np.float32)
    act, score = orc.densify_decide(projs, views, l1s, thr, 8.0, with_score=True)
    for i in range(len(g["pos"])):
        cand, rmax, best = False, 0.0, 0.0
        for pr, v, l1, th in zip(projs, views, l1s, thr):
            if pr["tiles_count"][i] == 0:
                continue
            rmax = max(rmax, float(pr["radius"][i]))
            m, r = pr["mean2d"][i], np.float32(pr["radius"][i])
            x0 = min(max(int(np.floor(m[0] / np.float32(16) - r / np.float32(16))), 0), v.tiles_x)
            x1 = min(max(int(np.ceil(m[0] / np.float32(16) + r / np.float32(16))), 0), v.tiles_x)
            y0 = min(max(int(np.floor(m[1] / np.float32(16) - r / np.float32(16))), 0), v.tiles_y)
            y1 = min(max(int(np.ceil(m[1] / np.float32(16) + r / np.float32(16))), 0), v.tiles_y)
            tiles = [ty * v.tiles_x + tx for ty in range(y0, y1) for tx in range(x0, x1)]
            cand = cand or any(l1[t] > th for t in tiles)
            best = max([best] + [float(np.float32(l1[t]) / np.float32(th)) for t in tiles])
        want = 0 if not cand else (2 if rmax > 8.0 else 1)
        assert act[i] == want, i
        # the normalised score: the worst tile over the threshold, > 1 for candidates
        assert score[i] == (np.float32(best) if cand else 0.0), i
        assert (score[i] > 1.0) == cand, i
    assert (act == 1).any() and (act == 2).any() and (act == 0).any()


def test_densify_cap_order():
    """The growth cap keeps the max_new highest scores, ties in index order
    (gaussian-renderer.ts:3297-3306: Array.sort is stable), by brute force."""
    import oracle as orc

    rng = np.random.default_rng(9)
    n = 4000
    act = rng.choice([0, 1, 2], n, p=[0.5, 0.3, 0.2]).astype(np.uint8)
    score = np.where(act > 0, rng.integers(1, 40, n) / 8.0, 0.0).astype(np.float32)  # many ties
    cand = [i for i in range(n) if act[i]]
    for max_new in (0, 1, 7, 400, len(cand) - 1, len(cand), len(cand) + 5):
        got = orc.densify_cap(act, score, max_new)
        keep = set(sorted(cand, key=lambda i: (-float(score[i]), i))[:max_new]) if 0 < max_new < len(cand) else set(cand)
        for i in range(n):
            assert got[i] == (act[i] if i in keep else 0), (max_new, i)
    assert int(np.count_nonzero(orc.densify_cap(act, score, max(1, n // 10)))) == max(1, n // 10)


def _rule(**kw):
    r = {"grad_threshold": 0.125, "grad_abs_threshold": 0.25, "size_threshold": 1.0, "min_visibility": 25.0,
         "importance_threshold": 5.0, "pixel_scale": 8.0}
    r.update(kw)
    return r


def test_densify_decide_grad_boundaries(orc):
    """performDensification's candidate tests (gaussian-renderer.ts:3262-3290)
    at their boundaries. With gradThreshold 0.125, gradAbsThreshold 0.25 and
    pixel scale 8 the thresholds are exactly 1 and 2 px; size threshold 1
    (exp(0) = 1 exactly); 2 samples, so importance must exceed 10."""
    rows = [  # (sum |g2d|, sum |gconic|, visible steps), log-scales, importance -> action, score
        ((1e6, 1e6, 24.0), (0.0, 0.0, 0.0), 100, 0, 0.0),    # fewer than 25 visible steps
        ((25.0, 0.0, 25.0), (0.0, -1.0, -1.0), 11, 1, 1.0),  # mean 2-D gradient = threshold: clone
        ((25.0, 0.0, 25.0), (0.0, -1.0, -1.0), 10, 0, 0.0),  # importance 10 / 2 = 5, not above 5
        ((24.99, 1e6, 25.0), (0.0, 0.0, 0.0), 50, 0, 0.0),   # just below; the conic test is for large splats
        ((0.0, 100.0, 50.0), (0.01, -1.0, -1.0), 50, 2, 1.0),  # large, mean conic gradient = 2: split
        ((1000.0, 49.0, 25.0), (-1.0, 0.01, -1.0), 50, 0, 0.0),  # large with a small conic gradient
        ((120.0, 0.0, 40.0), (-2.0, -2.0, -2.0), 50, 1, 3.0),  # clone, score 3
        ((0.0, 200.0, 25.0), (-1.0, -1.0, 0.5), 50, 2, 4.0),   # split, score 8 / 2
    ]
    acc = np.array([[r[0][0], 0.0, r[0][1], r[0][2]] for r in rows], np.float32)
    ls = np.array([r[1] for r in rows], np.float32)
    imp = np.array([r[2] for r in rows], np.uint32)
    act, score = orc.densify_decide_grad(acc, ls, imp, 2, _rule(), with_score=True)
    assert act.tolist() == [r[3] for r in rows]
    assert score.tolist() == [r[4] for r in rows]
    # visibility is at least 1 whatever the setting; a zero pixel threshold
    # ranks by the mean itself
    acc2 = np.array([[5.0, 0.0, 0.0, 0.5], [5.0, 0.0, 0.0, 1.0]], np.float32)
    act2, score2 = orc.densify_decide_grad(acc2, np.zeros((2, 3), np.float32), np.full(2, 99, np.uint32), 1,
                                           _rule(min_visibility=0.0, grad_threshold=0.0), with_score=True)
    assert act2.tolist() == [0, 1] and score2.tolist() == [0.0, 5.0]
    # one sample counts as max(1, samples): importance 6 > 5
    act3 = orc.densify_decide_grad(acc[1:2], ls[1:2], np.array([6], np.uint32), 0, _rule())
    assert act3.tolist() == [1]


def test_densify_grad_cap_keeps_best_scores(orc):
    """The growth cap on the gradient rule's scores: highest first, equal
    scores in index order."""
    rows = [(25.0, 25.0), (25.0, 25.0), (75.0, 25.0), (50.0, 25.0)]  # scores 1, 1, 3, 2
    acc = np.array([[a, 0.0, 0.0, d] for a, d in rows], np.float32)
    act, score = orc.densify_decide_grad(acc, np.full((4, 3), -2.0, np.float32), np.full(4, 50, np.uint32), 2,
                                         _rule(), with_score=True)
    assert act.tolist() == [1, 1, 1, 1] and score.tolist() == [1.0, 1.0, 3.0, 2.0]
    assert orc.densify_cap(act, score, 2).tolist() == [0, 0, 1, 1]
    assert orc.densify_cap(act, score, 3).tolist() == [1, 0, 1, 1]
    assert orc.densify_cap(act, score, 0).tolist() == [1, 1, 1, 1]
