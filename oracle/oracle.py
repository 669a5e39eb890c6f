"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, where it is the checker. The product library
(skyfall-viewer_amd/libsv.so) never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcCamera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("view_inv", C.c_float * 16), ("proj", C.c_float * 16),
                ("proj_inv", C.c_float * 16), ("viewport", C.c_float * 2), ("focal", C.c_float * 2)]


class OrcView(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("tile_size", C.c_uint32),
                ("tiles_x", C.c_uint32), ("tiles_y", C.c_uint32), ("sh_deg", C.c_uint32),
                ("near_plane", C.c_float), ("far_plane", C.c_float), ("radius_clip", C.c_float),
                ("eps2d", C.c_float), ("bg", C.c_float * 3), ("geom_flags", C.c_uint32)]


class OrcParams(C.Structure):
    _fields_ = [(k, C.POINTER(C.c_float)) for k in ("pos", "logscale", "quat", "logit", "sh")]


class OrcGrads(C.Structure):
    _fields_ = [(k, C.POINTER(C.c_float)) for k in ("pos", "logscale", "quat", "logit", "sh", "geom")]


class OrcStepOut(C.Structure):
    _fields_ = [("proj", C.c_void_p), ("keys", C.POINTER(C.c_uint32)), ("vals", C.POINTER(C.c_uint32)),
                ("cap_isect", C.c_size_t), ("offsets", C.POINTER(C.c_uint32)),
                ("color", C.POINTER(C.c_float)), ("alpha", C.POINTER(C.c_float)),
                ("last_ids", C.POINTER(C.c_uint32)), ("residual", C.POINTER(C.c_float)),
                ("tile_grads_fixed", C.POINTER(C.c_int32))]


PROJ_DTYPE = np.dtype([("mean2d", "<f4", 2), ("radius", "<f4"), ("depth", "<f4"), ("conic", "<f4", 3),
                       ("tiles_count", "<u4"), ("position_world", "<f4", 3), ("opacity_logit", "<f4")])
assert PROJ_DTYPE.itemsize == 48

WF_REFERENCE_ORDER, WF_SKIP_098 = 0x1, 0x2
NO_LAST = 0xFFFFFFFF

_lib = None
_fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
_u32p = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))  # noqa: E731
_i32p = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
_dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_expf.restype = C.c_float
        L.orc_expf.argtypes = [C.c_float]
        L.orc_f16_round.restype = C.c_float
        L.orc_f16_round.argtypes = [C.c_float]
        L.orc_encode_key.restype = C.c_uint32
        L.orc_encode_key.argtypes = [C.c_uint32, C.c_uint32, C.c_float]
        L.orc_count_isects.restype = C.c_uint64
        L.orc_render_and_grad.restype = C.c_int64
        L.orc_forward_loss_f64.restype = C.c_double
        L.orc_forward_loss_f64_only.restype = C.c_double
        L.orc_compact_rows.restype = C.c_size_t
        _lib = L
    return _lib


def view_params(w, h, sh_deg=3, tile=16, bg=(0.0, 0.0, 0.0), near=0.2, far=1e4, radius_clip=512.0,
                eps2d=0.3, geom_flags=0) -> OrcView:
    v = OrcView()
    v.geom_flags = geom_flags
    v.width, v.height, v.tile_size = w, h, tile
    v.tiles_x, v.tiles_y = (w + tile - 1) // tile, (h + tile - 1) // tile
    v.sh_deg = sh_deg
    v.near_plane, v.far_plane, v.radius_clip, v.eps2d = near, far, radius_clip, eps2d
    for i in range(3):
        v.bg[i] = bg[i]
    return v


def camera(cam_struct) -> OrcCamera:
    """Copy any ctypes camera with the CameraUniforms fields."""
    o = OrcCamera()
    C.memmove(C.byref(o), C.byref(cam_struct), C.sizeof(OrcCamera))
    return o


class Params:
    """Keeps numpy arrays alive behind an OrcParams struct."""

    def __init__(self, g: dict):
        self.a = {k: np.ascontiguousarray(g[k], dtype=np.float32) for k in ("pos", "logscale", "quat", "logit", "sh")}
        self.n = self.a["pos"].shape[0]
        self.s = OrcParams(*[_fp(self.a[k]) for k in ("pos", "logscale", "quat", "logit", "sh")])


def zero_grads(n: int) -> dict:
    return {"pos": np.zeros((n, 3), np.float32), "logscale": np.zeros((n, 3), np.float32),
            "quat": np.zeros((n, 4), np.float32), "logit": np.zeros((n,), np.float32),
            "sh": np.zeros((n, 48), np.float32), "geom": np.zeros((n, 5), np.float32)}


def grads_struct(g: dict) -> OrcGrads:
    return OrcGrads(*[_fp(g[k]) for k in ("pos", "logscale", "quat", "logit", "sh", "geom")])


def expf(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().orc_expf_v(C.c_size_t(x.size), _fp(x), _fp(y))
    return y


def exp_alpha(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().orc_exp_alpha_v(C.c_size_t(x.size), _fp(x), _fp(y))
    return y


def f16_round(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().orc_f16_round_v(C.c_size_t(x.size), _fp(x), _fp(y))
    return y


def sh_color(dirs: np.ndarray, sh: np.ndarray, deg: int) -> np.ndarray:
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    s = np.ascontiguousarray(sh, np.float32).reshape(-1, 48)
    o = np.empty((d.shape[0], 3), np.float32)
    lib().orc_sh_color_v(C.c_size_t(d.shape[0]), _fp(d), _fp(s), C.c_uint32(deg), _fp(o))
    return o


def project(g: dict, cam, view: OrcView) -> np.ndarray:
    p = Params(g)
    out = np.zeros(p.n, PROJ_DTYPE)
    oc = camera(cam)
    lib().orc_project(C.c_size_t(p.n), C.byref(p.s), C.byref(oc), C.byref(view), out.ctypes.data_as(C.c_void_p))
    return out


def emit(proj: np.ndarray, view: OrcView):
    m = int(lib().orc_count_isects(C.c_size_t(proj.shape[0]), proj.ctypes.data_as(C.c_void_p)))
    keys = np.zeros(max(m, 1), np.uint32)
    vals = np.zeros(max(m, 1), np.uint32)
    lib().orc_emit(C.c_size_t(proj.shape[0]), proj.ctypes.data_as(C.c_void_p), C.byref(view), _u32p(keys), _u32p(vals))
    return keys[:m].copy(), vals[:m].copy()


def sort_pairs(keys: np.ndarray, vals: np.ndarray):
    k = np.array(keys, np.uint32, copy=True)
    v = np.array(vals, np.uint32, copy=True)
    lib().orc_sort_pairs(C.c_size_t(k.size), _u32p(k), _u32p(v))
    return k, v


def sort_pairs_radix(keys: np.ndarray, vals: np.ndarray):
    """The parallel LSD radix sort (same output as sort_pairs)."""
    k = np.array(keys, np.uint32, copy=True)
    v = np.array(vals, np.uint32, copy=True)
    lib().orc_sort_pairs_radix(C.c_size_t(k.size), _u32p(k), _u32p(v))
    return k, v


def tile_offsets(keys: np.ndarray, n_tiles: int) -> np.ndarray:
    k = np.ascontiguousarray(keys, np.uint32)
    o = np.zeros(n_tiles + 1, np.uint32)
    lib().orc_tile_offsets(C.c_size_t(k.size), _u32p(k), C.c_uint32(n_tiles), _u32p(o))
    return o


def loss(color, target, w_l2, w_l1, w_ssim=0.0):
    c = np.ascontiguousarray(color, np.float32)
    t = np.ascontiguousarray(target, np.float32)
    h, w = c.shape[:2]
    r = np.empty_like(c)
    lib().orc_loss(_fp(c), _fp(t), C.c_uint32(w), C.c_uint32(h), C.c_float(w_l2), C.c_float(w_l1),
                   C.c_float(w_ssim), _fp(r))
    return r


def adam_bias(t: int, beta1: float, beta2: float):
    a, b = C.c_float(), C.c_float()
    lib().orc_adam_bias(C.c_uint32(t), C.c_float(beta1), C.c_float(beta2), C.byref(a), C.byref(b))
    return a.value, b.value


def adam(param, m1, m2, grad, lr, beta1, beta2, eps, bc1, bc2s):
    """In place on float32 arrays."""
    lib().orc_adam(C.c_size_t(param.size), _fp(param), _fp(m1), _fp(m2), _fp(np.ascontiguousarray(grad, np.float32)),
                   C.c_float(lr), C.c_float(beta1), C.c_float(beta2), C.c_float(eps), C.c_float(bc1),
                   C.c_float(bc2s))


def render_and_grad(g: dict, cam, view: OrcView, target=None, w_l2=0.8, w_l1=0.2, w_ssim=0.0,
                    wavefront=False, wf_flags=0, cap_isect=None, grads=None) -> dict:
    """Project, sort, forward; with a target also loss, backward and the
    geometry backward. Returns every intermediate as numpy arrays. With
    `grads` (a zero_grads dict) the gradients are added into it in place,
    after its per-view grad_geom planes are cleared, as a training step
    accumulates views."""
    p = Params(g)
    n = p.n
    proj = project(g, cam, view)
    m = int(lib().orc_count_isects(C.c_size_t(n), proj.ctypes.data_as(C.c_void_p)))
    cap = max(m, 1) if cap_isect is None else cap_isect
    W, H = view.width, view.height
    P = W * H
    nt = view.tiles_x * view.tiles_y
    out = {
        "proj": np.zeros(n, PROJ_DTYPE), "keys": np.zeros(cap, np.uint32), "vals": np.zeros(cap, np.uint32),
        "offsets": np.zeros(nt + 1, np.uint32), "color": np.zeros((H, W, 3), np.float32),
        "alpha": np.zeros((H, W), np.float32), "last_ids": np.zeros((H, W), np.uint32),
        "residual": np.zeros((H, W, 3), np.float32),
    }
    fixed = np.zeros((cap, 54), np.int32) if wavefront else None
    so = OrcStepOut(out["proj"].ctypes.data_as(C.c_void_p), _u32p(out["keys"]), _u32p(out["vals"]), cap,
                    _u32p(out["offsets"]), _fp(out["color"]), _fp(out["alpha"]), _u32p(out["last_ids"]),
                    _fp(out["residual"]), _i32p(fixed) if fixed is not None else None)
    if grads is None:
        grads = zero_grads(n)
    else:
        grads["geom"][...] = 0.0
    gs = grads_struct(grads)
    oc = camera(cam)
    tgt = None if target is None else np.ascontiguousarray(target, np.float32)
    r = lib().orc_render_and_grad(C.c_size_t(n), C.byref(p.s), C.byref(oc), C.byref(view),
                                  _fp(tgt) if tgt is not None else None, C.c_float(w_l2), C.c_float(w_l1),
                                  C.c_float(w_ssim), 1 if wavefront else 0, C.c_uint32(wf_flags), C.byref(so),
                                  C.byref(gs) if tgt is not None else None)
    if r < 0:
        raise RuntimeError("oracle capacity")
    out["n_isect"] = int(r)
    out["keys"] = out["keys"][:r]
    out["vals"] = out["vals"][:r]
    if fixed is not None:
        out["tile_grads_fixed"] = fixed[:r]
    if tgt is not None:
        out["grads"] = grads
    return out


def forward_loss_f64(g: dict, cam, view: OrcView, target, w_l2, w_l1, want_color=False):
    n = g["pos"].shape[0]
    arrs = [np.ascontiguousarray(g[k], np.float64) for k in ("pos", "logscale", "quat", "logit", "sh")]
    t = np.ascontiguousarray(target, np.float32)
    col = np.zeros((view.height, view.width, 3), np.float64) if want_color else None
    oc = camera(cam)
    val = lib().orc_forward_loss_f64(C.c_size_t(n), *[_dp(a) for a in arrs], C.byref(oc), C.byref(view), _fp(t),
                                     C.c_double(w_l2), C.c_double(w_l1), _dp(col) if col is not None else None)
    return (val, col) if want_color else val


def forward_loss_f64_only(g: dict, cam, view: OrcView, target, w_l2, w_l1, only: int):
    """The FP64 loss over the pixels of Gaussian `only`'s tile rectangle (for
    central differences of that Gaussian's SH coefficients)."""
    n = g["pos"].shape[0]
    arrs = [np.ascontiguousarray(g[k], np.float64) for k in ("pos", "logscale", "quat", "logit", "sh")]
    t = np.ascontiguousarray(target, np.float32)
    oc = camera(cam)
    return lib().orc_forward_loss_f64_only(C.c_size_t(n), *[_dp(a) for a in arrs], C.byref(oc), C.byref(view),
                                           _fp(t), C.c_double(w_l2), C.c_double(w_l1), C.c_size_t(only))


def compact_rows(src: np.ndarray, keep: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(src, np.float32)
    stride = int(np.prod(s.shape[1:])) if s.ndim > 1 else 1
    k = np.ascontiguousarray(keep, np.uint8)
    d = np.zeros_like(s)
    n = lib().orc_compact_rows(C.c_size_t(s.shape[0]), C.c_size_t(stride), _fp(s), k.ctypes.data_as(C.POINTER(C.c_uint8)),
                               _fp(d))
    return d[:n]


PARAM_KEYS = ("pos", "logscale", "quat", "logit", "sh")
PARAM_WIDTH = {"pos": 3, "logscale": 3, "quat": 4, "logit": 1, "sh": 48}


def to_rows(g: dict) -> np.ndarray:
    """Parameters (or moments) as rows of 59 floats: pos, log-scale, quat, logit, SH."""
    n = g["pos"].shape[0]
    return np.ascontiguousarray(np.concatenate([np.asarray(g[k], np.float32).reshape(n, -1) for k in PARAM_KEYS], 1))


def from_rows(rows: np.ndarray) -> dict:
    out, c = {}, 0
    for k in PARAM_KEYS:
        w = PARAM_WIDTH[k]
        out[k] = np.ascontiguousarray(rows[:, c:c + w] if w > 1 else rows[:, c])
        c += w
    return out


def prune_keep(logit: np.ndarray, min_opacity: float) -> np.ndarray:
    lg = np.ascontiguousarray(logit, np.float32)
    keep = np.zeros(lg.size, np.uint8)
    lib().orc_prune_keep(C.c_size_t(lg.size), _fp(lg), C.c_float(min_opacity), keep.ctypes.data_as(C.POINTER(C.c_uint8)))
    return keep


def accumulate_importance(proj: np.ndarray, stats4: np.ndarray, decay: float) -> np.ndarray:
    """accumulate_importance (importance-accum.wgsl:19-40) of one view's
    projections into a copy of stats4 (n, 4): importance_sum, visibility,
    max screen radius, pad."""
    st = np.array(stats4, np.float32, copy=True).reshape(-1, 4)
    pr = np.ascontiguousarray(proj)
    lib().orc_accumulate_importance(C.c_size_t(st.shape[0]), pr.ctypes.data_as(C.c_void_p), C.c_float(decay),
                                    _fp(st))
    return st


def accumulate_grads(geom: np.ndarray, accum4: np.ndarray) -> np.ndarray:
    """accumulate_grads (densify-accum.wgsl:19-46) of one view's geometric
    gradients geom (n, 5: conic a, b, c, mean x, y) into a copy of accum4
    (n, 4): |dL/dmean2d|, 0, |dL/dconic|, count."""
    acc = np.array(accum4, np.float32, copy=True).reshape(-1, 4)
    gg = np.ascontiguousarray(geom, np.float32).reshape(-1, 5)
    lib().orc_accumulate_grads(C.c_size_t(acc.shape[0]), _fp(gg), _fp(acc))
    return acc


def knn_init(pos: np.ndarray, extent: float):
    """(density, logscale) per point; pos (n, 3)."""
    p = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    n = p.shape[0]
    dens = np.zeros(n, np.float32)
    ls = np.zeros(n, np.float32)
    lib().orc_knn_init(C.c_size_t(n), _fp(p), C.c_float(extent), _fp(dens), _fp(ls))
    return dens, ls


def multiview_scores(g: dict, cams, views, targets, n_views=None):
    """generateMultiViewScores over the given views in order
    (gaussian-renderer.ts:2754-2781): per view render, score
    (orc_metric_view), then importance += counts and pruning += counts *
    mean (a double sum stored to float32). Returns (importance u32,
    pruning f32, views scored)."""
    n = g["pos"].shape[0]
    imp = np.zeros(n, np.uint32)
    pru = np.zeros(n, np.float32)
    scored = 0
    for cam, view, tgt in zip(cams, views, targets):
        r = render_and_grad(g, cam, view)
        counts = np.zeros(n, np.uint32)
        mean = C.c_double(0.0)
        t = np.ascontiguousarray(tgt, np.float32)
        ok = lib().orc_metric_view(C.byref(view), C.c_size_t(n), r["proj"].ctypes.data_as(C.c_void_p),
                                   _u32p(r["offsets"]), _u32p(r["vals"]), _u32p(r["last_ids"]),
                                   _fp(r["color"]), _fp(t), _u32p(counts), C.byref(mean))
        if not ok:
            continue
        scored += 1
        imp += counts
        pru = (pru.astype(np.float64) + counts.astype(np.float64) * mean.value).astype(np.float32)
    return imp, pru, scored


def densify(g: dict, m1: dict, m2: dict, actions: np.ndarray, seed: int):
    """Returns (params, m1, m2) dicts after clone / split."""
    r, a1, a2 = to_rows(g), to_rows(m1), to_rows(m2)
    act = np.ascontiguousarray(actions, np.uint8)
    n = r.shape[0]
    n_out = int(n + np.count_nonzero(act == 1) + np.count_nonzero(act == 2))
    o, o1, o2 = (np.zeros((max(n_out, 1), 59), np.float32) for _ in range(3))
    lib().orc_densify.restype = C.c_size_t
    w = lib().orc_densify(C.c_size_t(n), _fp(r), _fp(a1), _fp(a2), act.ctypes.data_as(C.POINTER(C.c_uint8)),
                          C.c_uint64(seed), _fp(o), _fp(o1), _fp(o2))
    assert w == n_out
    return from_rows(o[:w]), from_rows(o1[:w]), from_rows(o2[:w])


def tile_l1(view: OrcView, color: np.ndarray, target: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(color, np.float32)
    t = np.ascontiguousarray(target, np.float32)
    out = np.zeros(view.tiles_x * view.tiles_y, np.float32)
    lib().orc_tile_l1(C.byref(view), _fp(c), _fp(t), _fp(out))
    return out


def densify_decide(projs, views, tile_l1s, thr, split_px: float, with_score: bool = False):
    """Actions per Gaussian (0 none, 1 clone, 2 split); with_score also
    returns the candidates' normalised scores."""
    nv = len(projs)
    n = projs[0].shape[0]
    parr = (C.c_void_p * nv)(*[p.ctypes.data_as(C.c_void_p) for p in projs])
    varr = (OrcView * nv)(*views)
    tl = [np.ascontiguousarray(t, np.float32) for t in tile_l1s]
    tarr = (C.POINTER(C.c_float) * nv)(*[_fp(t) for t in tl])
    th = np.ascontiguousarray(thr, np.float32)
    act = np.zeros(n, np.uint8)
    score = np.zeros(n, np.float32)
    lib().orc_densify_decide(C.c_size_t(n), C.c_size_t(nv), parr, varr, tarr, _fp(th), C.c_float(split_px),
                             act.ctypes.data_as(C.POINTER(C.c_uint8)), _fp(score))
    return (act, score) if with_score else act


def densify_decide_grad(accum4: np.ndarray, logscale: np.ndarray, importance: np.ndarray, n_samples: int,
                        rule: dict, with_score: bool = False):
    """The reference trainer's densify candidates (orc_densify_decide_grad);
    rule keys: grad_threshold, grad_abs_threshold, size_threshold,
    min_visibility, importance_threshold, pixel_scale."""
    acc = np.ascontiguousarray(accum4, np.float32).reshape(-1, 4)
    ls = np.ascontiguousarray(logscale, np.float32).reshape(-1, 3)
    imp = np.ascontiguousarray(importance, np.uint32)
    n = acc.shape[0]
    r = (C.c_double * 6)(*[float(rule[k]) for k in DENSIFY_RULE_KEYS])
    act = np.zeros(n, np.uint8)
    score = np.zeros(n, np.float64)
    lib().orc_densify_decide_grad(C.c_size_t(n), _fp(acc), _fp(ls), _u32p(imp), C.c_uint32(n_samples), r,
                                  act.ctypes.data_as(C.POINTER(C.c_uint8)),
                                  score.ctypes.data_as(C.POINTER(C.c_double)))
    return (act, score) if with_score else act


DENSIFY_RULE_KEYS = ("grad_threshold", "grad_abs_threshold", "size_threshold", "min_visibility",
                     "importance_threshold", "pixel_scale")


def densify_cap(actions: np.ndarray, score: np.ndarray, max_new: int) -> np.ndarray:
    """performDensification's growth cap (gaussian-renderer.ts:3297-3306):
    the candidates sorted by score, highest first, with a stable sort of the
    index-ordered list (ties in index order); only the first max_new keep
    their action."""
    act = np.array(actions, np.uint8, copy=True)
    cand = np.flatnonzero(act)
    if max_new <= 0 or cand.size <= max_new:
        return act
    order = cand[np.argsort(-np.asarray(score, np.float64)[cand], kind="stable")]
    act[order[max_new:]] = 0
    return act


def percentile_lower(x: np.ndarray, p: float) -> float:
    """The value at index floor(p (n - 1)) of the sorted values."""
    s = np.sort(np.asarray(x, np.float32))
    return float(s[int(np.floor(p * (s.size - 1)))])


def position_lr(lr_pos: float, t: int) -> float:
    """computePositionLearningRate (gaussian-renderer.ts:284-298) in float64,
    stored as float32: log-linear decay to 1.6e-6 over 30000 steps with a
    sine warm-up from 0.01 x over the first 3000."""
    import math

    if not lr_pos > 0.0:
        return float(np.float32(lr_pos))
    lr_final, delay_mult, max_steps = 0.0000016, 0.01, 30000.0
    step = max(float(t), 0.0)
    tt = min(1.0, step / max_steps)
    loglerp = math.exp(math.log(float(np.float32(lr_pos))) * (1 - tt) + math.log(lr_final) * tt)
    phase = min(1.0, step / max(1.0, round(0.1 * max_steps)))
    delay = delay_mult + (1.0 - delay_mult) * math.sin(0.5 * math.pi * phase)
    return float(np.float32(loglerp * delay))


def adam_update(params: dict, m1: dict, m2: dict, grads: dict, t: int, lrs: dict, beta1: float, beta2: float,
                eps: float, sh_update_interval: int = 1, pos_lr_schedule: bool = False):
    """run_adam_updates (gaussian-renderer.ts:2455-2545) in place: adam_step
    (adam.wgsl:18-38) per tensor with its learning rate, t = step index; the
    position rate follows the schedule when asked and is skipped when not
    positive; SH only when t % sh_update_interval == 0 (:2482-2484)."""
    bc1, bc2 = adam_bias(t, beta1, beta2)
    for k in PARAM_KEYS:
        lr = lrs[k]
        if k == "pos":
            lr = position_lr(lr, t) if pos_lr_schedule else lr
            if not np.float32(lr) > 0:
                continue
        if k == "sh" and sh_update_interval and t % sh_update_interval != 0:
            continue
        adam(params[k], m1[k], m2[k], grads[k], lr, beta1, beta2, eps, bc1, bc2)


def train(g: dict, cams, views, targets, schedule, lrs: dict, beta1=0.9, beta2=0.999, eps=1e-15, w_l2=0.8,
          w_l1=0.2, w_ssim=0.0, sh_update_interval=1, pos_lr_schedule=False, wavefront=False, wf_flags=0,
          t0=0, state=None):
    """trainStep (gaussian-renderer.ts:3638-3660) for every entry of
    `schedule` (a list of view-index lists): zero the gradients, render,
    loss and backward each view in order (gradients summed; grad_geom per
    view), then one Adam step. Returns (params, m1, m2, per-step list of
    (params copy, grads)). `state` = (params, m1, m2) continues a run."""
    if state is None:
        p = {k: np.array(g[k], np.float32, copy=True) for k in PARAM_KEYS}
        m1 = {k: np.zeros_like(p[k]) for k in PARAM_KEYS}
        m2 = {k: np.zeros_like(p[k]) for k in PARAM_KEYS}
    else:
        p, m1, m2 = ({k: np.array(d[k], np.float32, copy=True) for k in PARAM_KEYS} for d in state)
    hist = []
    n = p["pos"].shape[0]
    for i, vs in enumerate(schedule):
        gr = zero_grads(n)
        for v in vs:
            render_and_grad(p, cams[v], views[v], target=targets[v], w_l2=w_l2, w_l1=w_l1, w_ssim=w_ssim,
                            wavefront=wavefront, wf_flags=wf_flags, grads=gr)
        adam_update(p, m1, m2, gr, t0 + i + 1, lrs, beta1, beta2, eps, sh_update_interval, pos_lr_schedule)
        hist.append(({k: p[k].copy() for k in PARAM_KEYS}, gr))
    return p, m1, m2, hist
