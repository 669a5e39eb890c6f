"""Pins of the CPU oracle against what fixes each function independently of
the oracle's own code: libm / numpy, closed forms, invariants, brute force
and float64 autograd (CPU only).

The reference ships no golden vectors or known-answer tests for this path,
and its WGSL cannot run here, so these checks are what makes the oracle
trustworthy. Each test names the oracle function it pins and is chosen so a
plausible slip (a dropped term, a wrong sign, index or transposition) fails.
"""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, close_rel

NO_LAST = 0xFFFFFFFF


def _ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, np.int64(-2**31) - a, a)
    b = np.where(b < 0, np.int64(-2**31) - b, b)
    return np.abs(a - b)


# ---------------------------------------------------------------- orc_expf
def test_expf_against_libm(orc):
    rng = np.random.default_rng(0)
    x = np.concatenate([np.linspace(-103.0, 88.7, 200001, dtype=np.float32),
                        rng.uniform(-20, 5, 100000).astype(np.float32),
                        np.array([0.0, -0.0, 1.0, -1.0, 88.72, -87.3, -87.4, -100.0], np.float32)])
    got = orc.expf(x)
    ref = np.exp(x.astype(np.float64)).astype(np.float32)
    normal = ref >= np.float32(np.finfo(np.float32).tiny)
    assert _ulp_diff(got[normal], ref[normal]).max() <= 2
    # gradual underflow: within one unit of the subnormal spacing
    sub = ~normal
    assert np.abs(got[sub].astype(np.float64) - ref[sub].astype(np.float64)).max() <= 2 * 2.0**-149
    assert orc.expf(np.array([0.0], np.float32))[0] == 1.0
    sp = orc.expf(np.array([100.0, -200.0, np.inf, -np.inf, np.nan], np.float32))
    assert sp[0] == np.inf and sp[1] == 0.0 and sp[2] == np.inf and sp[3] == 0.0 and np.isnan(sp[4])


# ----------------------------------------------------------- orc_exp_alpha
def test_exp_alpha_within_wgsl_bound(orc):
    """The splat-weight exp against float64 exp: within 3 + 2|x| ulp (WGSL's
    accuracy requirement for exp) on [-80, 1/2], exhaustively over every
    float in [-1, 1/2] and on 2^21 random points beyond; exact at 0 to
    within its polynomial's 1 ulp; monotone over a dense grid."""
    rng = np.random.default_rng(5)
    lo = np.float32(-1.0).view(np.uint32)
    neg = np.arange(np.uint32(0x80000000), lo + 1, 37, dtype=np.uint32).view(np.float32)  # [-1, -0] strided
    pos = np.arange(0, np.float32(0.5).view(np.uint32) + 1, 37, dtype=np.uint32).view(np.float32)
    x = np.concatenate([neg, pos, rng.uniform(-80, 0.5, 1 << 21).astype(np.float32),
                        np.float32([-80.0, -79.999, -0.0, 0.0, 0.5, -6.0, -5.5])])
    got = orc.exp_alpha(x).astype(np.float64)
    ref64 = np.exp(x.astype(np.float64))
    ulp = np.abs(got - ref64) / np.spacing(ref64.astype(np.float32)).astype(np.float64)
    assert (ulp <= 3.0 + 2.0 * np.abs(x.astype(np.float64))).all(), float((ulp / (3 + 2 * np.abs(x))).max())
    assert ulp[np.abs(x) < 1].max() <= 2.5 and ulp[np.abs(x) < 6].max() <= 7.5
    assert orc.exp_alpha(np.float32([0.0, -0.0]))[0] == 1.0
    g = np.linspace(-80, 0.5, 1_000_001, dtype=np.float32)
    e = orc.exp_alpha(g)
    assert (np.diff(e.astype(np.float64)) >= 0).all()


# ------------------------------------------------------------ orc_f16_round
def test_f16_round_matches_numpy_rne(orc):
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2**32, 400000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    # plus values around the half subnormal / normal / overflow boundaries and exact ties
    edge = np.array([6.1035156e-05, 6.0975552e-05, 5.9604645e-08, 2.9802322e-08, 2.9802326e-08,
                     65504.0, 65519.996, 65520.0, 1.0 + 2.0**-11, 1.0 + 3 * 2.0**-11, -0.0, 0.0],
                    np.float32)
    x = np.concatenate([x, edge, rng.uniform(-4, 4, 100000).astype(np.float32)])
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).astype(np.float32)
    got = orc.f16_round(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


# ----------------------------------------------------- orc_sh_color (basis)
def _sh_basis_from_oracle(orc, dirs, deg=3):
    """Basis values Y_k(dir) read through orc_sh_color with one unit
    coefficient at a time (scaled so the +0.5 / clamp at 0 never binds)."""
    n = dirs.shape[0]
    out = np.zeros((n, 16))
    for k in range(16):
        sh = np.zeros((n, 48), np.float32)
        sh[:, 3 * k] = 0.25
        rgb = orc.sh_color(dirs, sh, deg).astype(np.float64)
        out[:, k] = (rgb[:, 0] - 0.5) / 0.25
    return out


def test_sh_basis_orthonormal_with_reference_signs(orc):
    # Gauss-Legendre in cos(theta) x uniform phi: exact for the degree <= 6
    # products of the basis
    zt, wt = np.polynomial.legendre.leggauss(16)
    nphi = 32
    phi = 2 * np.pi * np.arange(nphi) / nphi
    z = np.repeat(zt, nphi)
    w = np.repeat(wt, nphi) * (2 * np.pi / nphi)
    r = np.sqrt(1 - z * z)
    p = np.tile(phi, 16)
    dirs = np.stack([r * np.cos(p), r * np.sin(p), z], 1).astype(np.float32)
    Y = _sh_basis_from_oracle(orc, dirs)
    gram = (Y * w[:, None]).T @ Y
    assert np.abs(gram - np.eye(16)).max() < 2e-5
    # Condon-Shortley signs of the reference basis (common.wgsl:1-18,
    # training-forward-tiles.wgsl:65-105): -C1 y, +C1 z, -C1 x
    c1 = math.sqrt(3.0 / (4.0 * math.pi))
    ax = _sh_basis_from_oracle(orc, np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], np.float32))
    assert abs(ax[0, 1] + c1) < 1e-5 and abs(ax[1, 2] - c1) < 1e-5 and abs(ax[2, 3] + c1) < 1e-5
    assert abs(ax[0, 0] - 0.5 / math.sqrt(math.pi)) < 1e-6
    # degree truncation: bands above sh_deg contribute nothing
    d1 = _sh_basis_from_oracle(orc, dirs[:50], deg=1)
    assert np.abs(d1[:, 4:]).max() == 0.0 and np.abs(d1[:, :4] - Y[:50, :4]).max() < 1e-6


# ------------------------------------------------------------ orc_project
def _one(pos, logscale, quat=(0, 0, 0, 1), logit=2.0):
    n = len(pos)
    sh = np.zeros((n, 48), np.float32)
    sh[:, 0:3] = 0.5
    return {"pos": np.asarray(pos, np.float32).reshape(n, 3),
            "logscale": np.asarray(logscale, np.float32).reshape(n, 3),
            "quat": np.tile(np.asarray(quat, np.float32), (n, 1)),
            "logit": np.full(n, logit, np.float32), "sh": sh}


@pytest.mark.parametrize("sigma", [0.002, 0.01, 0.05, 0.2])
def test_projection_on_axis_isotropic_closed_form(sv, orc, sigma):
    W, H = 256, 192
    cam = sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], W, H, 60.0)
    view = orc.view_params(W, H, sh_deg=3)
    g = _one([[0, 0, 0]], [[math.log(sigma)] * 3], quat=(0.3, -0.2, 0.5, 0.7))
    pr = orc.project(g, cam, view)[0]
    f = float(cam.focal[0])
    assert abs(f - (H / 2) / math.tan(math.radians(30))) < 1e-3
    s2 = sigma * sigma * f * f / 9.0 + 0.3
    assert abs(pr["depth"] - 3.0) < 1e-6
    assert abs(pr["conic"][0] * s2 - 1) < 1e-5 and abs(pr["conic"][2] * s2 - 1) < 1e-5
    assert abs(pr["conic"][1]) < 1e-6 * pr["conic"][0]
    assert pr["radius"] == math.ceil(3 * math.sqrt(s2))
    assert pr["mean2d"][0] == W / 2 and pr["mean2d"][1] == H / 2
    r = pr["radius"]
    x0, x1 = max(0, math.floor((W / 2 - r) / 16)), min(W // 16, math.ceil((W / 2 + r) / 16))
    y0, y1 = max(0, math.floor((H / 2 - r) / 16)), min(H // 16, math.ceil((H / 2 + r) / 16))
    assert pr["tiles_count"] == (x1 - x0) * (y1 - y0)


def test_projection_culled_means_no_tiles(sv, orc):
    cam = sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], 128, 128, 60.0)
    view = orc.view_params(128, 128, sh_deg=3)
    # behind the camera, far outside the frustum, nearer than the near plane
    g = _one([[0, 0, 5], [40, 0, 0], [0, 0, 2.9]], [[-3.0] * 3] * 3)
    pr = orc.project(g, cam, view)
    assert np.all(pr["tiles_count"] == 0) and np.all(pr["radius"] == 0)


# ---------------------------------------------- orc_sort_pairs, orc_tile_offsets
@pytest.mark.parametrize("monotone_vals", [True, False])
def test_sort_is_stable_sort_by_key(orc, monotone_vals):
    rng = np.random.default_rng(2)
    n = 50000
    keys = (rng.integers(0, 300, n).astype(np.uint32) << 16) | rng.integers(0, 4, n).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32) if monotone_vals else rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    k, v = orc.sort_pairs(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])


@pytest.mark.parametrize("bits", [7, 16, 28, 32])
@pytest.mark.parametrize("monotone_vals", [True, False])
def test_radix_sort_is_stable_sort_by_key(orc, bits, monotone_vals):
    rng = np.random.default_rng(bits)
    n = 200003
    keys = rng.integers(0, 2 ** bits, n, dtype=np.uint64).astype(np.uint32)
    keys[::7] = keys[3]  # many ties
    vals = np.arange(n, dtype=np.uint32) if monotone_vals else rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    k, v = orc.sort_pairs_radix(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])
    k2, v2 = orc.sort_pairs(keys, vals)
    assert np.array_equal(k, k2) and np.array_equal(v, v2)
    assert orc.sort_pairs_radix(keys[:1], vals[:1])[0].tolist() == keys[:1].tolist()


def test_tile_offsets_are_lower_bounds(orc):
    rng = np.random.default_rng(3)
    nt = 700
    tiles = np.sort(rng.choice(np.arange(0, nt, 3), 20000))  # two of every three tiles empty
    keys = (tiles.astype(np.uint32) << 16) | rng.integers(0, 60000, tiles.size).astype(np.uint32)
    keys.sort()
    off = orc.tile_offsets(keys, nt)
    assert np.array_equal(off, np.searchsorted(keys >> 16, np.arange(nt + 1), "left").astype(np.uint32))
    assert orc.tile_offsets(np.zeros(0, np.uint32), 5).tolist() == [0] * 6


# ------------------------------------------------------------- orc_emit
def test_emit_invariants(sv, orc):
    g = sv.scene_gaussians(0, 4096, 1.0)
    cam = sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], 256, 256, 60.0)
    view = orc.view_params(256, 256, sh_deg=3)
    pr = orc.project(g, cam, view)
    keys, vals = orc.emit(pr, view)
    counts = pr["tiles_count"].astype(np.int64)
    assert keys.size == counts.sum()
    assert np.array_equal(np.bincount(vals, minlength=len(pr)), counts)
    # emit order: Gaussian index, then ty, then tx
    assert np.all(np.diff(vals.astype(np.int64)) >= 0)
    tiles = (keys >> 16) & 0xFFF
    dq = keys & 0xFFFF
    starts = np.concatenate([[0], np.cumsum(counts)])
    ts = np.float32(16)
    for gi in np.flatnonzero(counts)[:500]:
        t = tiles[starts[gi]:starts[gi + 1]]
        tx, ty = t % 16, t // 16
        assert np.all(np.diff(ty * 1000 + tx) > 0)
        m = pr["mean2d"][gi].astype(np.float32)
        r = np.float32(pr["radius"][gi])
        x0 = min(max(int(np.floor(m[0] / ts - r / ts)), 0), 16)
        x1 = min(max(int(np.ceil(m[0] / ts + r / ts)), 0), 16)
        y0 = min(max(int(np.floor(m[1] / ts - r / ts)), 0), 16)
        y1 = min(max(int(np.ceil(m[1] / ts + r / ts)), 0), 16)
        want = [yy * 16 + xx for yy in range(y0, y1) for xx in range(x0, x1)]
        assert t.tolist() == want
        assert np.all(dq[starts[gi]:starts[gi + 1]] == min(int(pr["depth"][gi]), 65535))


# ------------------------------------------------------------ orc_forward
def _view_dirs(cam, pos):
    V = np.array(cam.view, np.float64).reshape(4, 4).T
    pc = pos.astype(np.float64) @ V[:3, :3].T + V[:3, 3]
    d = -pc
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def test_forward_matches_brute_force_compositing(sv, orc):
    W, H = 96, 80
    g = sv.scene_gaussians(7, 2048, 1.0)
    g["logscale"] = g["logscale"] + np.float32(1.0)
    cam = sv.scene_camera([0.3, -0.2, 3.0], [0.0, 0.0, 0.0], W, H, 60.0)
    bg = (0.1, 0.2, 0.3)
    view = orc.view_params(W, H, sh_deg=3, bg=bg)
    r = orc.render_and_grad(g, cam, view)
    pr = r["proj"]
    rgb = np.maximum(orc.sh_color(_view_dirs(cam, g["pos"]), g["sh"], 3).astype(np.float64), 0.0)
    op = 1.0 / (1.0 + np.exp(-g["logit"].astype(np.float64)))
    off, vals = r["offsets"], r["vals"]
    tx_n = (W + 15) // 16
    for t in range(len(off) - 1):
        tx, ty = t % tx_n, t // tx_n
        ids = vals[off[t]:off[t + 1]]
        for py in range(ty * 16, min(ty * 16 + 16, H)):
            for px in range(tx * 16, min(tx * 16 + 16, W)):
                C = np.zeros(3)
                Acc = 0.0
                last = NO_LAST
                for j, gi in enumerate(ids):
                    last = off[t] + j
                    dx, dy = px + 0.5 - pr["mean2d"][gi, 0], py + 0.5 - pr["mean2d"][gi, 1]
                    a, b, c = pr["conic"][gi]
                    s = 0.5 * (a * dx * dx + c * dy * dy) + b * dx * dy
                    if s < 0:
                        continue
                    al = min(op[gi] * math.exp(-s), 0.999)
                    if al < 1.0 / 255.0:
                        continue
                    C += rgb[gi] * al * (1 - Acc)
                    Acc += al * (1 - Acc)
                    if 1 - Acc < 1e-4:
                        break
                out = C + (1 - Acc) * np.array(bg)
                assert np.abs(out - r["color"][py, px]).max() < 1e-5, (px, py)
                assert abs(Acc - r["alpha"][py, px]) < 1e-5
                if len(ids) == 0:
                    assert r["last_ids"][py, px] == NO_LAST
                else:
                    assert r["last_ids"][py, px] == last


def test_forward_special_cases(sv, orc):
    W, H = 64, 64
    bg = (0.25, 0.5, 0.75)
    cam = sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], W, H, 60.0)
    view = orc.view_params(W, H, sh_deg=0, bg=bg)
    # one opaque splat on the axis: every other tile is empty
    lg = math.log(0.9 / 0.1)
    g = _one([[0, 0, 0]], [[math.log(0.01)] * 3], logit=lg)
    r = orc.render_and_grad(g, cam, view)
    pr = r["proj"][0]
    a, b, c = pr["conic"]
    dx = dy = 0.5  # pixel (32, 32) against the mean (32, 32)
    al = min(0.9 * math.exp(-(0.5 * (a * dx * dx + c * dy * dy) + b * dx * dy)), 0.999)
    rgb = 0.5 * 0.28209479177387814 * np.ones(3) + 0.5
    want = rgb * al + (1 - al) * np.array(bg)
    assert np.abs(r["color"][32, 32] - want).max() < 1e-6
    assert r["last_ids"][32, 32] == r["offsets"][(32 // 16) * 4 + 32 // 16]
    # empty tiles: background, zero alpha, no last index
    assert np.abs(r["color"][0, 0] - np.array(bg)).max() == 0.0
    assert r["alpha"][0, 0] == 0.0 and r["last_ids"][0, 0] == NO_LAST
    # termination after adding: three stacked nearly opaque splats; the pixel
    # stops at the second (1 - A = (1 - 0.992)^2 < 1e-4) and never visits the third
    lg = math.log(0.9955 / 0.0045)
    g3 = _one([[0, 0, 0], [0, 0, 0.01], [0, 0, 0.02]], [[math.log(0.5)] * 3] * 3, logit=lg)
    r3 = orc.render_and_grad(g3, cam, view)
    t = (32 // 16) * 4 + 32 // 16
    assert r3["offsets"][t + 1] - r3["offsets"][t] == 3
    assert r3["last_ids"][32, 32] == r3["offsets"][t] + 1
    a1 = 1.0 - r3["alpha"][32, 32]
    assert a1 < 1e-4


# --------------------------------------------------------------- orc_loss
def test_loss_residual_is_gradient_of_l2_l1(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    c = rng.uniform(0, 1, (20, 24, 3)).astype(np.float32)
    t = rng.uniform(0, 1, (20, 24, 3)).astype(np.float32)
    t[0, 0] = c[0, 0]  # d = 0: sign(0) = 0
    res = orc.loss(c, t, 0.8, 0.2)
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    d = ct - torch.tensor(t, dtype=torch.float64)
    L = (0.8 * 0.5 * d * d + 0.2 * d.abs()).sum()
    L.backward()
    assert np.abs(res - ct.grad.numpy()).max() < 1e-6
    assert np.all(res[0, 0] == 0.0)


# ---------------------------------------------------------- orc_adam
def test_adam_first_step_closed_form_and_n_steps(orc):
    rng = np.random.default_rng(5)
    n = 10000
    g0 = rng.normal(0, 1, n).astype(np.float32) * np.float32(10.0) ** rng.integers(-4, 2, n).astype(np.float32)
    p = rng.normal(0, 1, n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-15
    bc1, bc2 = orc.adam_bias(1, b1, b2)
    q = p.copy()
    orc.adam(q, m, v, g0, lr, b1, b2, eps, bc1, bc2)
    delta = (q.astype(np.float64) - p) / lr
    want = -g0.astype(np.float64) / (np.abs(g0.astype(np.float64)) + eps)
    assert np.abs(delta - want).max() < 1e-3  # f32 rounding of p - lr*... over |p| <= 5
    # 20 steps against a float64 textbook Adam
    P, M, V = p.astype(np.float64), np.zeros(n), np.zeros(n)
    q, m, v = p.copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)
    for t in range(1, 21):
        grad = (g0 * np.float32(math.cos(t))).astype(np.float32)
        bc1, bc2 = orc.adam_bias(t, b1, b2)
        orc.adam(q, m, v, grad, lr, b1, b2, eps, bc1, bc2)
        M = b1 * M + (1 - b1) * grad
        V = b2 * V + (1 - b2) * grad.astype(np.float64) ** 2
        P = P - lr * (M / (1 - b1 ** t)) / (np.sqrt(V / (1 - b2 ** t)) + eps)
    assert np.abs(q - P).max() < 2e-5
    assert close_rel(m, M, 1e-5) < 1e-5


# ------------------------------------------------------ orc_compact_rows
def test_compaction_is_boolean_indexing(orc):
    rng = np.random.default_rng(6)
    src = rng.normal(size=(1000, 59)).astype(np.float32)
    keep = (rng.uniform(size=1000) < 0.7).astype(np.uint8)
    assert np.array_equal(orc.compact_rows(src, keep), src[keep.astype(bool)])


def test_termination_threshold_is_the_reference_test():
    """The forward ends a pixel when 1 - A < 1e-4 (training-forward-tiles.wgsl,
    TRANSMITTANCE_THRESHOLD); the library tests A > SV_A_DONE instead (one
    compare). Over every float A in [1/2, 2] the two tests agree, and below
    1/2 both are false (1 - A rounds to at least 1/2 there; checked on a
    sample). The constant is read from the device header."""
    import re

    hdr = open(os.path.join(ROOT, "skyfall-viewer_amd", "csrc", "sv_device.hpp")).read()
    m = re.search(r"#define SV_A_DONE (0x[0-9a-fp.+-]+)f", hdr)
    assert m, "SV_A_DONE not found"
    thr = np.float32(float.fromhex(m.group(1)))
    eps = np.float32(1e-4)
    lo, hi = np.float32(0.5).view(np.uint32), np.float32(2.0).view(np.uint32)
    a = np.arange(lo, hi + 1, dtype=np.uint32).view(np.float32)
    assert np.array_equal((np.float32(1.0) - a) < eps, a > thr)
    small = np.random.default_rng(9).uniform(0, 0.5, 1 << 20).astype(np.float32)
    assert not ((np.float32(1.0) - small) < eps).any() and not (small > thr).any()
