"""Pins of the oracle's SSIM loss term (orc_loss with w_ssim > 0), the
D-SSIM residual of training-loss.wgsl:33-84:

    SSIM = (2 mu_x mu_y + C1)(2 s_xy + C2) / ((mu_x^2 + mu_y^2 + C1)(s_x^2 + s_y^2 + C2))
    residual += w_ssim * (1 - SSIM) / 2 * (pred - target)

over a 5x5 window whose coordinates are clamped to the image (samplePred
clamps; the target sampler clamps to the edge), C1 = 0.01^2, C2 = 0.03^2,
population (1/25) moments. The checks below do not retype that loop: they
evaluate the window moments in closed form from how many window columns (or
rows) fall on each side of an edge, in float64.

- uniform images: every window variance is 0, so
  SSIM = (2ab + C1) / (a^2 + b^2 + C1) exactly (C2 cancels);
- a step edge in the prediction (and, separately, in the target too): with
  f = (window columns left of the edge) / 5, mu = f p1 + (1 - f) p2,
  s^2 = f (1 - f) (p1 - p2)^2 and s_xy = f (1 - f) (p1 - p2)(t1 - t2),
  which also pins the clamping at the image border (a pixel in column 0
  counts column 0 three times).

A plausible mistake fails one of them: C1 and C2 swapped (uniform case), a
3x3 or zero-padded window (edge near the border), sample instead of
population variance, a dropped factor 1/2 or a flipped (pred - target).
"""
import numpy as np
import pytest

C1 = 0.01 ** 2
C2 = 0.03 ** 2


def _ssim_term(mx, my, sx2, sy2, sxy, p, t):
    s = (2 * mx * my + C1) * (2 * sxy + C2) / ((mx * mx + my * my + C1) * (sx2 + sy2 + C2))
    return 0.5 * (1.0 - s) * (p - t)


@pytest.mark.parametrize("a,b", [(0.25, 0.75), (0.5, 0.5), (1.0, 0.0), (0.0, 0.125), (0.3, 0.71), (0.9, 0.2)])
def test_uniform_images(orc, a, b):
    h, w = 9, 11
    pred = np.full((h, w, 3), a, np.float32)
    targ = np.full((h, w, 3), b, np.float32)
    got = orc.loss(pred, targ, 0.0, 0.0, 1.0).astype(np.float64)
    a64, b64 = float(np.float32(a)), float(np.float32(b))
    ssim = (2 * a64 * b64 + C1) / (a64 * a64 + b64 * b64 + C1)
    want = 0.5 * (1.0 - ssim) * (a64 - b64)
    assert np.allclose(got, want, rtol=2e-6, atol=1e-9), (got.ravel()[:3], want)


def _window_fraction(c, edge, n):
    """Fraction of the 5 clamped window positions around c that lie below edge."""
    pos = np.clip(np.arange(c - 2, c + 3), 0, n - 1)
    return np.count_nonzero(pos < edge) / 5.0


@pytest.mark.parametrize("axis", [1, 0])
@pytest.mark.parametrize("edge", [1, 2, 5, 9])
@pytest.mark.parametrize("target_edge", [False, True])
def test_step_edge(orc, axis, edge, target_edge):
    h, w = 10, 12
    p1, p2 = np.float32([0.8, 0.1, 0.45]), np.float32([0.2, 0.6, 0.05])
    t1, t2 = np.float32([0.5, 0.3, 0.9]), (np.float32([0.1, 0.7, 0.25]) if target_edge else np.float32([0.5, 0.3, 0.9]))
    idx = np.arange(w if axis == 1 else h)
    side = (idx < edge)[None, :, None] if axis == 1 else (idx < edge)[:, None, None]
    pred = np.where(side, p1, p2).astype(np.float32) * np.ones((h, w, 3), np.float32)
    targ = np.where(side, t1, t2).astype(np.float32) * np.ones((h, w, 3), np.float32)
    got = orc.loss(pred, targ, 0.0, 0.0, 1.0).astype(np.float64)
    n = w if axis == 1 else h
    want = np.zeros((h, w, 3))
    P1, P2, T1, T2 = (x.astype(np.float64) for x in (p1, p2, t1, t2))
    for y in range(h):
        for x in range(w):
            c = x if axis == 1 else y
            f = _window_fraction(c, edge, n)
            mx, my = f * P1 + (1 - f) * P2, f * T1 + (1 - f) * T2
            v = f * (1 - f)
            p = P1 if c < edge else P2
            t = T1 if c < edge else T2
            want[y, x] = _ssim_term(mx, my, v * (P1 - P2) ** 2, v * (T1 - T2) ** 2, v * (P1 - P2) * (T1 - T2), p, t)
    err = np.abs(got - want)
    assert err.max() <= 1e-6 + 1e-5 * np.abs(want).max(), float(err.max())
    # the window really straddles the edge somewhere (the variances matter)
    assert np.abs(want).max() > 0


def test_ssim_adds_to_l1_l2(orc):
    """The weights combine linearly: w_l2 d + w_l1 sign(d) + w_ssim * term."""
    rng = np.random.default_rng(3)
    pred = rng.random((16, 20, 3), dtype=np.float32)
    targ = rng.random((16, 20, 3), dtype=np.float32)
    term = orc.loss(pred, targ, 0.0, 0.0, 1.0).astype(np.float64)
    d = pred.astype(np.float64) - targ
    got = orc.loss(pred, targ, 0.3, 0.5, 0.2).astype(np.float64)
    want = 0.3 * d + 0.5 * np.sign(d) + 0.2 * term
    assert np.abs(got - want).max() <= 1e-6
