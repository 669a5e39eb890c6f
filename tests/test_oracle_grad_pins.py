"""Pins of the oracle's gradients (CPU only):

  orc_backward_legacy    float64 torch autograd of front-to-back compositing
                         over the oracle's own tile lists, termination and
                         skip decisions (training-backward-tiles.wgsl:250-397)
  orc_backward_wavefront equals the legacy gradients to the 2^-16 fixed-point
                         truncation where every list fits one 64-entry batch
                         and no alpha exceeds 0.98 (backward-tiles-local.wgsl)
  orc_backward_reduce    integer brute force of the per-intersection sums
  orc_backward_geom      float64 torch autograd of the covariance / mean model
                         the shader differentiates (training-backward-geom.wgsl)
  SH / logit gradients   central finite differences of the FP64 CPU forward
                         (the logit and every one of the 48 SH coefficients
                         of 256 Gaussians)
  geometry gradients     finite-difference gap printed, not enforced: the
                         reference differentiates a model that differs from its
                         forward (transposed view in the position chain, the
                         conic-b factor 2, no quaternion normalisation Jacobian)
"""
import math
import multiprocessing
import os

import numpy as np
import pytest

from conftest import close_rel

torch = pytest.importorskip("torch")
F64 = torch.float64
NO_LAST = 0xFFFFFFFF


def _scene(sv, orc, n, w, h, seed=0, eye=(0.0, 0.0, 3.0), logscale_shift=0.0, logit_clip=None):
    g = sv.scene_gaussians(seed, n, 1.0)
    if logscale_shift:
        g["logscale"] = (g["logscale"] + np.float32(logscale_shift)).astype(np.float32)
    if logit_clip is not None:
        g["logit"] = np.clip(g["logit"], -logit_clip, logit_clip).astype(np.float32)
    cam = sv.scene_camera(list(eye), [0.0, 0.0, 0.0], w, h, 60.0)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(seed + 1, g["pos"], 0.02)
    view = orc.view_params(w, h, sh_deg=3)
    target = orc.render_and_grad(tg, cam, view)["color"]
    return g, cam, view, target


def _view_dirs(cam, pos):
    V = np.array(cam.view, np.float64).reshape(4, 4).T
    pc = pos.astype(np.float64) @ V[:3, :3].T + V[:3, 3]
    d = -pc
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def _basis(orc, dirs):
    n = dirs.shape[0]
    out = np.zeros((n, 16))
    for k in range(16):
        sh = np.zeros((n, 48), np.float32)
        sh[:, 3 * k] = 0.25
        out[:, k] = (orc.sh_color(dirs, sh, 3)[:, 0].astype(np.float64) - 0.5) / 0.25
    return out


def _autograd_compositing(orc, g, cam, view, r, t_from_stored_alpha=True):
    """d(sum residual * C)/d(logit, sh, conic, mean2d) in float64.

    The reference's backward does not keep transmittance: it starts each
    pixel from T = 1 - A with A the stored float32 final alpha and divides
    by (1 - alpha) on the way back, so every T it uses is the exact one
    times r = (1 - A_f32) / T_exact. With t_from_stored_alpha the pixel's
    upstream gradient is scaled by that (constant) r, which is what the
    reference computes; without it the result is the exact derivative."""
    W, H = view.width, view.height
    n = g["pos"].shape[0]
    pr = r["proj"]
    B = torch.tensor(_basis(orc, _view_dirs(cam, g["pos"])), dtype=F64)
    mean = torch.tensor(pr["mean2d"], dtype=F64, requires_grad=True)
    conic = torch.tensor(pr["conic"], dtype=F64, requires_grad=True)
    logit = torch.tensor(g["logit"], dtype=F64, requires_grad=True)
    sh = torch.tensor(g["sh"].reshape(n, 16, 3), dtype=F64, requires_grad=True)
    rgb = (B[:, :, None] * sh).sum(1) + 0.5
    assert float(rgb.detach().min()) > 0.0, "colour clamp would be active"
    op = torch.sigmoid(logit)
    res = torch.tensor(r["residual"], dtype=F64)
    off, vals, last = r["offsets"], r["vals"].astype(np.int64), r["last_ids"]
    bg = torch.tensor([view.bg[0], view.bg[1], view.bg[2]], dtype=F64)
    tx_n = view.tiles_x
    total = torch.zeros((), dtype=F64)
    for t in range(len(off) - 1):
        s0, s1 = int(off[t]), int(off[t + 1])
        if s1 <= s0:
            continue
        tx, ty = t % tx_n, t // tx_n
        ys, xs = np.meshgrid(np.arange(ty * 16, min(ty * 16 + 16, H)), np.arange(tx * 16, min(tx * 16 + 16, W)),
                             indexing="ij")
        ys, xs = ys.ravel(), xs.ravel()
        ids = torch.tensor(vals[s0:s1])
        px = torch.tensor(xs + 0.5, dtype=F64)[:, None]
        py = torch.tensor(ys + 0.5, dtype=F64)[:, None]
        dx = px - mean[ids, 0][None]
        dy = py - mean[ids, 1][None]
        a, b, c = conic[ids, 0][None], conic[ids, 1][None], conic[ids, 2][None]
        s = 0.5 * (a * dx * dx + c * dy * dy) + b * dx * dy
        al = torch.clamp(op[ids][None] * torch.exp(-s), max=0.999)
        lp = last[ys, xs].astype(np.int64)
        lrel = np.where(lp == NO_LAST, -1, lp - s0)
        j = np.arange(s1 - s0)[None, :]
        with torch.no_grad():
            mask = (s >= 0) & (al >= 1.0 / 255.0) & torch.tensor(j <= lrel[:, None])
        ae = torch.where(mask, al, torch.zeros_like(al))
        T = torch.cumprod(1.0 - ae, 1)
        Tx = torch.cat([torch.ones_like(T[:, :1]), T[:, :-1]], 1)
        C = ((ae * Tx)[:, :, None] * rgb[ids][None]).sum(1) + T[:, -1:] * bg[None]
        up = res[ys, xs]
        if t_from_stored_alpha:
            with torch.no_grad():
                t32 = torch.tensor((np.float32(1.0) - r["alpha"][ys, xs]).astype(np.float64), dtype=F64)
                ratio = torch.where(T[:, -1] > 0, t32 / T[:, -1], torch.ones_like(t32))
            up = up * ratio[:, None]
        total = total + (up * C).sum()
    total.backward()
    geom = torch.cat([conic.grad, mean.grad], 1).numpy()
    return {"logit": logit.grad.numpy(), "sh": sh.grad.reshape(n, 48).numpy(), "geom": geom}


def test_legacy_backward_matches_autograd(sv, orc):
    g, cam, view, target = _scene(sv, orc, 4096, 64, 64)
    r = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    ag = _autograd_compositing(orc, g, cam, view, r)
    exact = _autograd_compositing(orc, g, cam, view, r, t_from_stored_alpha=False)
    for k in ("logit", "sh", "geom"):
        e = close_rel(r["grads"][k], ag[k], 1e-4)
        ex = close_rel(r["grads"][k], exact[k], 1e-4)
        print(f"[pin] legacy backward {k} vs autograd: {e:.2e} (vs the exact derivative {ex:.2e}: "
              "transmittance recovered from the stored float32 alpha)")
        assert e < 2e-5, k
        assert ex < 1e-3, k


def _wf(orc, g, cam, view, target, flags):
    return orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2, wavefront=True, wf_flags=flags)


def test_wavefront_fixed_point_equals_legacy_where_a11_a12_inactive(sv, orc):
    g, cam, view, target = _scene(sv, orc, 200, 128, 128, seed=11, logit_clip=2.0)
    leg = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    assert np.diff(leg["offsets"].astype(np.int64)).max() <= 64, "a tile list spans two batches"
    ref = _wf(orc, g, cam, view, target, 1 | 2)  # the reference kernel: batch order + 0.98 skip
    oc = _wf(orc, g, cam, view, target, 0)
    assert np.array_equal(ref["tile_grads_fixed"], oc["tile_grads_fixed"])
    # each pair truncates each slot by < 2^-16; a Gaussian gets <= 256 pairs per tile
    counts = leg["proj"]["tiles_count"].astype(np.float64)
    bound = 256.0 * counts * 2.0 ** -16 + 1e-6
    for k, sl in (("logit", None), ("sh", 48), ("geom", 5)):
        d = np.abs(ref["grads"][k].astype(np.float64) - leg["grads"][k])
        bb = bound if sl is None else bound[:, None]
        assert np.all(d <= bb), k
        assert close_rel(ref["grads"][k], leg["grads"][k], 1e-3) < 2e-3, k


def test_wavefront_batch_order_matters_when_lists_span_batches(sv, orc):
    g, cam, view, target = _scene(sv, orc, 4096, 64, 64)
    assert np.diff(orc.render_and_grad(g, cam, view)["offsets"].astype(np.int64)).max() > 64
    ref = _wf(orc, g, cam, view, target, 1)
    oc = _wf(orc, g, cam, view, target, 0)
    leg = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    assert not np.array_equal(ref["tile_grads_fixed"], oc["tile_grads_fixed"])
    e_oc = close_rel(oc["grads"]["logit"], leg["grads"]["logit"], 1e-3)
    e_ref = close_rel(ref["grads"]["logit"], leg["grads"]["logit"], 1e-3)
    print(f"[pin] logit grad vs legacy: order-correct batches {e_oc:.2e}, reference batch order {e_ref:.2e}")
    assert e_oc < 1e-3 < e_ref


def test_reduce_is_integer_brute_force(sv, orc):
    g, cam, view, target = _scene(sv, orc, 4096, 64, 64)
    wf = _wf(orc, g, cam, view, target, 0)
    ints = wf["tile_grads_fixed"].astype(np.float64) / 65536.0
    ints[np.abs(ints) <= 1e-10] = 0.0
    n = g["pos"].shape[0]
    logit = np.zeros(n)
    np.add.at(logit, wf["vals"], ints[:, 0])
    sh = np.zeros((n, 48))
    np.add.at(sh, wf["vals"], ints[:, 6:])
    geom = np.zeros((n, 5))
    np.add.at(geom, wf["vals"], ints[:, 1:6])
    assert close_rel(wf["grads"]["logit"], logit, 1e-6) < 1e-6
    assert close_rel(wf["grads"]["sh"], sh, 1e-6) < 1e-6
    assert close_rel(wf["grads"]["geom"], geom, 1e-6) < 1e-6


# ----------------------------------------------------- geometry backward
def _m3mul(A, B):
    # WGSL order: matrices stored m[col][row]; (A*B)[i][j] = sum_k A[k][j] B[i][k]
    return torch.einsum("nkj,nik->nij", A, B)


def _geom_model_grads(g, cam, gup, ysign=1.0):
    """VJP of the model training-backward-geom.wgsl differentiates, by
    autograd: t moves with the position through transpose(W) (W = the view
    rotation read row-wise), cov2D = T^T cov3D^T T with T = W J and the
    frustum-clamped J, conic = (c, -b, a) / det, mean = (fx tx, fy ty) / tz,
    sigma = exp(log-scale), R from the normalised quaternion; the upstream
    gradient of conic b enters twice (the shader's factor 2). ysign = -1
    models the mean's y as the forward places it, H - c_y - fy ty / tz."""
    n = g["pos"].shape[0]
    view = np.array(cam.view, np.float64)
    Vm = view.reshape(4, 4).T  # math [row][col]
    p0 = g["pos"].astype(np.float64)
    t0 = p0 @ Vm[:3, :3].T + Vm[:3, 3]
    pos = torch.tensor(p0, dtype=F64, requires_grad=True)
    ls = torch.tensor(g["logscale"].astype(np.float64), dtype=F64, requires_grad=True)
    q32 = g["quat"].astype(np.float32)
    qn = q32 / np.sqrt((q32 * q32).sum(1, keepdims=True))
    q = torch.tensor(qn.astype(np.float64), dtype=F64, requires_grad=True)
    Wm = torch.zeros((n, 3, 3), dtype=F64)
    for c in range(3):
        for r_ in range(3):
            Wm[:, c, r_] = float(view[r_ * 4 + c])
    # t = t0 + W (pos - p0) in the shader's convention: dL/dpos = transpose(W) dL/dt
    dp = pos - torch.tensor(p0, dtype=F64)
    t = torch.tensor(t0, dtype=F64) + torch.einsum("nkj,nk->nj", Wm, dp)
    fx, fy = float(cam.focal[0]), float(cam.focal[1])
    limx = 1.3 * float(cam.viewport[0]) / (2 * fx)
    limy = 1.3 * float(cam.viewport[1]) / (2 * fy)
    tz = t[:, 2]
    tcx = torch.clamp(t[:, 0] / tz, -limx, limx) * tz
    tcy = torch.clamp(t[:, 1] / tz, -limy, limy) * tz
    z0 = torch.zeros_like(tz)
    J = torch.stack([torch.stack([fx / tz, z0, -fx * tcx / (tz * tz)], 1),
                     torch.stack([z0, fy / tz, -fy * tcy / (tz * tz)], 1),
                     torch.stack([z0, z0, z0], 1)], 1)
    T = _m3mul(Wm, J)
    sig = torch.exp(ls)
    S = torch.diag_embed(sig)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
        torch.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
        torch.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)], 1)
    M = _m3mul(S, R)
    cov3 = _m3mul(M.transpose(1, 2), M)
    cov2 = _m3mul(_m3mul(T.transpose(1, 2), cov3.transpose(1, 2)), T)
    a = cov2[:, 0, 0] + 0.3
    b = cov2[:, 0, 1]
    c = cov2[:, 1, 1] + 0.3
    det = a * c - b * b
    conic = torch.stack([c / det, -b / det, a / det], 1)
    mx = fx * t[:, 0] / tz
    my = ysign * fy * t[:, 1] / tz
    gu = torch.tensor(gup.astype(np.float64), dtype=F64)
    L = (conic[:, 0] * gu[:, 0] + 2.0 * conic[:, 1] * gu[:, 1] + conic[:, 2] * gu[:, 2]
         + mx * gu[:, 3] + my * gu[:, 4]).sum()
    L.backward()
    return {"pos": pos.grad.numpy(), "logscale": ls.grad.numpy(), "quat": q.grad.numpy()}


def test_geometry_backward_matches_autograd_of_shader_model(sv, orc):
    g, cam, view, target = _scene(sv, orc, 4096, 64, 64, eye=(0.4, -0.3, 3.0))
    r = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    gg = r["grads"]
    live = np.abs(gg["geom"]).sum(1) > 1e-10
    assert live.sum() > 1000
    ag = _geom_model_grads({k: v[live] for k, v in g.items()}, cam, gg["geom"][live])
    for k in ("pos", "logscale", "quat"):
        e = close_rel(gg[k][live], ag[k], 1e-4)
        print(f"[pin] geometry backward {k} vs autograd of the shader model: {e:.2e}")
        assert e < 1e-4, k
    assert np.all(gg["pos"][~live] == 0)


def test_geometry_mean_y_forward_matches_autograd(sv, orc):
    """geom_flags bit 0 (the library's SV_GEOM_MEAN_Y_FORWARD): the mean's y
    enters with the forward's sign; the covariance path is unchanged."""
    g, cam, _, target = _scene(sv, orc, 4096, 64, 64, eye=(0.4, -0.3, 3.0))
    view = orc.view_params(64, 64, sh_deg=3, geom_flags=1)
    r = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    gg = r["grads"]
    live = np.abs(gg["geom"]).sum(1) > 1e-10
    ag = _geom_model_grads({k: v[live] for k, v in g.items()}, cam, gg["geom"][live], ysign=-1.0)
    for k in ("pos", "logscale", "quat"):
        e = close_rel(gg[k][live], ag[k], 1e-4)
        print(f"[pin] geometry backward, mean y as the forward, {k} vs autograd: {e:.2e}")
        assert e < 1e-4, k
    r0 = orc.render_and_grad(g, cam, orc.view_params(64, 64, sh_deg=3), target=target, w_l2=0.8, w_l1=0.2)
    for k in ("logscale", "quat", "logit", "sh", "geom"):
        assert np.array_equal(r0["grads"][k], gg[k]), k


def test_position_gradient_direction_against_fd(sv, orc):
    """Central FD (h = 1e-4) of the FP64 forward for the positions of 40
    Gaussians: with the mean's y in the forward's convention the gradient
    points along the FD gradient (cosine >= 0.95); the shader as written
    (+fy, A8) is reported, its y sign being opposite to the forward's."""
    g, cam, _, target = _scene(sv, orc, 4096, 64, 64, eye=(0.4, -0.3, 3.0))
    cos = {}
    for flags in (0, 1):
        view = orc.view_params(64, 64, sh_deg=3, geom_flags=flags)
        r = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
        live = np.flatnonzero(r["proj"]["tiles_count"] > 0)
        pick = np.random.default_rng(3).choice(live, 40, replace=False)
        g64 = {k: v.astype(np.float64) for k, v in g.items()}
        a, fd = [], []
        for i in pick:
            for j in range(3):
                gp = {k: v.copy() for k, v in g64.items()}
                gm = {k: v.copy() for k, v in g64.items()}
                gp["pos"][i, j] += 1e-4
                gm["pos"][i, j] -= 1e-4
                fd.append((orc.forward_loss_f64(gp, cam, view, target, 0.8, 0.2)
                           - orc.forward_loss_f64(gm, cam, view, target, 0.8, 0.2)) / 2e-4)
                a.append(r["grads"]["pos"][i, j])
        a, fd = np.array(a), np.array(fd)
        cos[flags] = float(a @ fd / (np.linalg.norm(a) * np.linalg.norm(fd)))
    print(f"[pin] position gradient vs FD of the forward: cosine {cos[0]:+.3f} as the shader, "
          f"{cos[1]:+.3f} with the forward's y")
    assert cos[1] >= 0.95
    assert cos[0] < cos[1]


# ------------------------------------------------ finite differences (FP64)
def test_sh_and_logit_gradients_match_central_fd_of_fp64_forward(sv, orc):
    """Adjoint-check scene (configuration 2): 4096 Gaussians, 64x64, one view,
    SH degree 3. For 256 Gaussians picked with seed 3, the logit derivative
    (h = 1e-3) and the
    derivative along a random direction of the 48 SH coefficients are
    compared with the oracle's float32 gradient. A pair or sign switch
    inside [x - h, x + h] makes the central difference meaningless; such
    points are detected by disagreement between h and h/2 and left out
    (at most 10%)."""
    g, cam, view, target = _scene(sv, orc, 4096, 64, 64)
    r = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    gr = r["grads"]
    live = np.flatnonzero(r["proj"]["tiles_count"] > 0)
    pick = np.random.default_rng(3).choice(live, 256, replace=False)
    g64 = {k: v.astype(np.float64) for k, v in g.items()}
    rng = np.random.default_rng(4)

    def f(gd):
        return orc.forward_loss_f64(gd, cam, view, target, 0.8, 0.2)

    def cfd(key, i, direction, h):
        gp = {k: v.copy() for k, v in g64.items()}
        gm = {k: v.copy() for k, v in g64.items()}
        gp[key][i] = gp[key][i] + h * direction
        gm[key][i] = gm[key][i] - h * direction
        return (f(gp) - f(gm)) / (2 * h)

    rows = []
    for i in pick:
        d = rng.normal(size=48)
        d /= np.linalg.norm(d)
        for key, direction, h, ana in (("logit", 1.0, 1e-3, float(gr["logit"][i])),
                                       ("sh", d, 1e-3, float(gr["sh"][i] @ d))):
            a1 = cfd(key, i, direction, h)
            a2 = cfd(key, i, direction, h / 2)
            rows.append((key, ana, a1, a2))
    ana = np.array([x[1] for x in rows])
    f1 = np.array([x[2] for x in rows])
    f2 = np.array([x[3] for x in rows])
    scale = np.abs(ana).max()
    smooth = np.abs(f1 - f2) <= 1e-5 * scale
    err = np.abs(ana - f2)[smooth].max() / scale
    print(f"[pin] FD SH/logit: {smooth.sum()}/{len(rows)} smooth points, max err {err:.2e} of max|g| {scale:.3e}")
    assert smooth.mean() >= 0.9
    assert err < 1e-4


# state the worker processes of the per-coefficient check read (set by
# _fd_init in each worker; each worker perturbs its own copy in place)
_FD = {}


def _fd_init(g64, cam, view, target):
    import oracle

    _FD.update(orc=oracle, g64=g64, cam=cam, view=view, target=target)


def _fd_sh_gaussian(i):
    """Central differences of the FP64 forward's loss for each of the 48 SH
    coefficients of Gaussian i, at h = 1e-3 and h / 2. The loss is summed over
    the pixels of the Gaussian's tile rectangle only: the other pixels' terms
    do not depend on its colour and cancel in the difference."""
    orc, g64, cam, view, target = _FD["orc"], _FD["g64"], _FD["cam"], _FD["view"], _FD["target"]
    sh = g64["sh"]
    out = np.zeros((48, 2))
    for k in range(48):
        x0 = sh[i, k]
        for col, h in enumerate((1e-3, 5e-4)):
            sh[i, k] = x0 + h
            fp = orc.forward_loss_f64_only(g64, cam, view, target, 0.8, 0.2, i)
            sh[i, k] = x0 - h
            fm = orc.forward_loss_f64_only(g64, cam, view, target, 0.8, 0.2, i)
            sh[i, k] = x0
            out[k, col] = (fp - fm) / (2 * h)
    return out


def test_sh_gradients_match_central_fd_per_coefficient(sv, orc):
    """Every parameter's derivative, not a projection: for each of the 48 SH
    coefficients of the 256 Gaussians of the test above (configuration 2's
    scene, seed 3), the central difference (h = 1e-3) of the FP64 forward's
    loss against the oracle's float32 gradient. As above, points where h and
    h / 2 disagree (an L1 sign or colour clamp switching inside the step) are
    left out, at most 10 %. The loss of each evaluation is summed over the
    Gaussian's tile rectangle (checked against the whole image's difference
    on three coefficients); the 49k evaluations run in worker processes. They are spawned, not forked:
    this process has loaded the HIP runtime (the scene generator lives in
    libsv), and a forked child of it stalls."""
    g, cam, view, target = _scene(sv, orc, 4096, 64, 64)
    r = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    ana = r["grads"]["sh"]
    live = np.flatnonzero(r["proj"]["tiles_count"] > 0)
    pick = np.random.default_rng(3).choice(live, 256, replace=False)
    g64 = {k: v.astype(np.float64) for k, v in g.items()}
    # the rectangle-restricted loss differences equal the whole image's
    for i, k in ((int(pick[0]), 0), (int(pick[1]), 17), (int(pick[2]), 47)):
        d = {}
        for sign in (1, -1):
            gd = {key: v.copy() for key, v in g64.items()}
            gd["sh"][i, k] += sign * 1e-3
            d[sign] = (orc.forward_loss_f64(gd, cam, view, target, 0.8, 0.2),
                       orc.forward_loss_f64_only(gd, cam, view, target, 0.8, 0.2, i))
        full, part = d[1][0] - d[-1][0], d[1][1] - d[-1][1]
        assert abs(full - part) <= 1e-9 * max(abs(full), 1e-12) + 1e-13, (i, k, full, part)
        assert d[1][1] < d[1][0]  # a strict part of the image
    # one OpenMP thread per worker (the forward's sort is OpenMP-parallel;
    # eight workers of eight threads each would oversubscribe the cores)
    saved = os.environ.get("OMP_NUM_THREADS")
    os.environ["OMP_NUM_THREADS"] = "1"
    try:
        with multiprocessing.get_context("spawn").Pool(max(1, min(8, os.cpu_count() or 1)), initializer=_fd_init,
                                                       initargs=(g64, cam, view, target)) as pool:
            fd = np.stack(pool.map(_fd_sh_gaussian, [int(i) for i in pick], chunksize=4))
    finally:
        if saved is None:
            del os.environ["OMP_NUM_THREADS"]
        else:
            os.environ["OMP_NUM_THREADS"] = saved
    a = ana[pick].astype(np.float64)  # (256, 48)
    f1, f2 = fd[:, :, 0], fd[:, :, 1]
    scale = np.abs(a).max()
    smooth = np.abs(f1 - f2) <= 1e-5 * scale
    err = np.abs(a - f2)[smooth].max() / scale
    # every coefficient index is checked somewhere, and on most Gaussians
    per_coeff = smooth.mean(axis=0)
    print(f"[pin] FD per SH coefficient: {smooth.sum()}/{smooth.size} smooth points, max err {err:.2e} of "
          f"max|g| {scale:.3e}; fewest smooth points of one coefficient {per_coeff.min():.2f}")
    assert smooth.mean() >= 0.9
    assert per_coeff.min() >= 0.5
    assert err < 1e-4


def test_geometry_fd_gap_is_reported(sv, orc):
    """The reference's geometry gradients are matched to the reference (above);
    their gap to central differences (h = 1e-4) of the FP64 forward is
    printed, not enforced, for every position, log-scale and quaternion
    component of 256 Gaussians picked with seed 3 (configuration 2's scene,
    camera off the axis): the largest difference over the largest FD value
    and the cosine between the two gradient vectors, per parameter kind."""
    g, cam, view, target = _scene(sv, orc, 4096, 64, 64, eye=(0.4, -0.3, 3.0))
    r = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    gr = r["grads"]
    pick = np.random.default_rng(3).choice(np.flatnonzero(r["proj"]["tiles_count"] > 0), 256, replace=False)
    g64 = {k: v.astype(np.float64) for k, v in g.items()}
    gaps = {}
    for key, m in (("pos", 3), ("logscale", 3), ("quat", 4)):
        a, fd = [], []
        for i in pick:
            for j in range(m):
                gp = {k: v.copy() for k, v in g64.items()}
                gm = {k: v.copy() for k, v in g64.items()}
                gp[key][i, j] += 1e-4
                gm[key][i, j] -= 1e-4
                fd.append((orc.forward_loss_f64(gp, cam, view, target, 0.8, 0.2)
                           - orc.forward_loss_f64(gm, cam, view, target, 0.8, 0.2)) / 2e-4)
                a.append(gr[key][i, j])
        a, fd = np.array(a), np.array(fd)
        gaps[key] = (float(np.abs(a - fd).max() / max(np.abs(fd).max(), 1e-30)),
                     float(a @ fd / max(np.linalg.norm(a) * np.linalg.norm(fd), 1e-30)))
    print("[pin] geometry gradient gap to FD of the forward, 256 Gaussians (reported): "
          + ", ".join(f"{k} max {v[0]:.2e} cos {v[1]:+.3f}" for k, v in gaps.items()))
    assert all(math.isfinite(v[0]) and math.isfinite(v[1]) for v in gaps.values())