"""The rasterizers' culling bounds, restated in float32 numpy and checked
by brute force (the HIP kernels sv_raster.hip quad_mask / rect_min_power
and sv_internal.hpp tight_insets follow the same operations; the GPU parity
tests check the kernels themselves against the oracle).

- The quadrant test keeps every quadrant of a tile that holds a pixel the
  per-pixel tests take (power in [0, power_max] inside the alpha box), for
  splats of every shape: isotropic to axis ratio e^3 (conic condition up to
  ~400), all orientations, opacities from 0.005 to 1, means inside and
  outside the tile.
- The tight tile range contains every tile where a quadrant box test
  passes, including at the image border.
- The edge-minimum formula: on a quadrant that holds the mean the minimum
  is 0; otherwise it equals the brute-force minimum over a fine grid of
  the quadrant's boundary up to rounding.
"""
import numpy as np

f = np.float32


def _splats(rng, n, span):
    mx = rng.uniform(-span, 16 + span, n).astype(f)
    my = rng.uniform(-span, 16 + span, n).astype(f)
    s1 = np.exp(rng.uniform(-1, 3.5, n))
    s2 = s1 * np.exp(rng.uniform(-3, 0, n))
    th = rng.uniform(0, np.pi, n)
    c, s = np.cos(th), np.sin(th)
    cxx = c * c * s1 * s1 + s * s * s2 * s2 + 0.3
    cyy = s * s * s1 * s1 + c * c * s2 * s2 + 0.3
    cxy = c * s * (s1 * s1 - s2 * s2)
    det = cxx * cyy - cxy * cxy
    a, b, cc = (cyy / det).astype(f), (-cxy / det).astype(f), (cxx / det).astype(f)
    op = rng.uniform(0.005, 1.0, n)
    pmax = (np.log(255.0 * op) + 0.02).astype(f)
    qdet = a * cc - b * b
    ex = (np.sqrt(f(2) * pmax * cc / qdet) * f(1.01) + f(0.01)).astype(f)
    ey = (np.sqrt(f(2) * pmax * a / qdet) * f(1.01) + f(0.01)).astype(f)
    return mx, my, a, b, cc, pmax, ex, ey


def _rect_min_power(mx, my, a, b, c, x0, x1, y0, y1):
    half = f(0.5)
    inv_a, inv_c = f(1) / a, f(1) / c
    dx = np.minimum(np.maximum(mx, x0), x1) - mx
    dy = np.minimum(np.maximum(my, y0), y1) - my
    best = np.where((dx != 0) | (dy != 0), f(np.inf), f(0)).astype(f)
    s = np.zeros_like(mx)
    t = np.minimum(np.maximum(-b * dx * inv_c, y0 - my), y1 - my)
    p1 = half * (a * dx * dx + c * t * t) + b * dx * t
    best = np.where(dx != 0, np.minimum(best, p1), best)
    s = np.where(dx != 0, half * (a * dx * dx + c * t * t) + np.abs(b * dx * t), s)
    t = np.minimum(np.maximum(-b * dy * inv_a, x0 - mx), x1 - mx)
    p2 = half * (a * t * t + c * dy * dy) + b * t * dy
    best = np.where(dy != 0, np.minimum(best, p2), best)
    s = np.where(dy != 0, np.maximum(s, half * (a * t * t + c * dy * dy) + np.abs(b * t * dy)), s)
    return best, s


def test_quadrant_test_keeps_every_taking_pixel():
    rng = np.random.default_rng(7)
    n = 60000
    mx, my, a, b, c, pmax, ex, ey = _splats(rng, n, 40.0)
    lo, hi = [f(0.5), f(8.5)], [f(7.5), f(15.5)]
    kept_box = dropped = 0
    for j in range(4):
        x0, x1, y0, y1 = lo[j & 1], hi[j & 1], lo[j >> 1], hi[j >> 1]
        box = (np.abs(mx - np.minimum(np.maximum(mx, x0), x1)) <= ex) & \
              (np.abs(my - np.minimum(np.maximum(my, y0), y1)) <= ey)
        mn, s = _rect_min_power(mx, my, a, b, c, x0, x1, y0, y1)
        sbox = f(0.5) * (a * ex * ex + c * ey * ey) + np.abs(b) * ex * ey
        keep = box & ~(mn > pmax + f(2e-6) * (sbox + s))
        xs = np.arange(8, dtype=f) + x0
        ys = np.arange(8, dtype=f) + y0
        PX, PY = [v.ravel() for v in np.meshgrid(xs, ys)]
        DX = PX[None, :] - mx[:, None]
        DY = PY[None, :] - my[:, None]
        sg = f(0.5) * (a[:, None] * DX * DX + c[:, None] * DY * DY) + b[:, None] * DX * DY
        take = (np.abs(DX) <= ex[:, None]) & (np.abs(DY) <= ey[:, None]) & (sg >= 0) & (sg <= pmax[:, None])
        assert not (take.any(1) & ~keep).any(), f"quadrant {j} dropped a taking pixel"
        kept_box += int(box.sum())
        dropped += int((box & ~keep).sum())
    assert kept_box > 0 and dropped > 0.05 * kept_box  # the ellipse test does drop quadrants


def test_tight_range_contains_every_box_pass():
    rng = np.random.default_rng(11)
    n = 400000
    W = 1000  # 63 tiles, the last one 8 pixels wide
    m = rng.uniform(-200, W + 200, n).astype(f)
    e = np.exp(rng.uniform(-4, 6, n)).astype(f)
    lo = np.ceil(np.clip((m - e - f(16.5)) * f(0.0625), -1e6, 1e6)).astype(np.int64)
    hi = np.floor(np.clip((m + e + f(0.5)) * f(0.0625), -1e6, 1e6)).astype(np.int64) + 1
    for t in range((W + 15) // 16):
        passed = np.zeros(n, bool)
        for h in range(2):
            xs = t * 16 + 8 * h
            if xs >= W:
                continue
            x0 = f(xs) + f(0.5)
            x1 = f(min(xs + 7, W - 1)) + f(0.5)
            passed |= np.abs(m - np.minimum(np.maximum(m, x0), x1)) <= e
        assert not (passed & ((t < lo) | (t >= hi))).any(), t


def test_edge_minimum_matches_fine_grid():
    rng = np.random.default_rng(5)
    mx, my, a, b, c, *_ = _splats(rng, 3000, 30.0)
    x0, x1, y0, y1 = f(0.5), f(7.5), f(0.5), f(7.5)
    mn, _ = _rect_min_power(mx, my, a, b, c, x0, x1, y0, y1)
    inside = (mx >= x0) & (mx <= x1) & (my >= y0) & (my <= y1)
    assert np.all(mn[inside] == 0)
    g = np.linspace(0.5, 7.5, 701)
    edges = np.concatenate([np.stack([np.full_like(g, 0.5), g], 1), np.stack([np.full_like(g, 7.5), g], 1),
                            np.stack([g, np.full_like(g, 0.5)], 1), np.stack([g, np.full_like(g, 7.5)], 1)])
    a64, b64, c64 = a.astype(np.float64), b.astype(np.float64), c.astype(np.float64)
    dx = edges[None, :, 0] - mx[:, None].astype(np.float64)
    dy = edges[None, :, 1] - my[:, None].astype(np.float64)
    brute = (0.5 * (a64[:, None] * dx * dx + c64[:, None] * dy * dy) + b64[:, None] * dx * dy).min(1)
    out = ~inside
    # the closed form is the exact minimum; the grid's is at most a grid step's worth above it
    assert np.all(mn[out] <= brute[out] * (1 + 1e-5) + 1e-5)
    assert np.all(mn[out] >= brute[out] - 0.02 * (1 + brute[out]))
