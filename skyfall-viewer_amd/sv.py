"""ctypes binding of libsv.so (include/sv_api.h) with numpy buffers.

This is plumbing for Python callers (tests, benchmarks): every call goes
straight to the HIP library. There is no fallback: if libsv.so is missing or
no GPU is present, the call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsv.so")

SV_OK, SV_E_ARG, SV_E_CAPACITY, SV_E_HIP, SV_E_STATE, SV_E_NOMEM = range(6)
MODE_PLAIN, MODE_BATCHED, MODE_SEQUENTIAL = 0, 1, 2
WF_REFERENCE_ORDER, WF_SKIP_098 = 0x1, 0x2
GEOM_MEAN_Y_FORWARD = 0x1
NO_LAST = 0xFFFFFFFF


class Camera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("view_inv", C.c_float * 16), ("proj", C.c_float * 16),
                ("proj_inv", C.c_float * 16), ("viewport", C.c_float * 2), ("focal", C.c_float * 2)]


class Config(C.Structure):
    _fields_ = [
        ("tile_size", C.c_uint32), ("batch", C.c_uint32), ("raster_px", C.c_uint32),
        ("sh_deg", C.c_uint32), ("bg", C.c_float * 3),
        ("near_plane", C.c_float), ("far_plane", C.c_float), ("radius_clip", C.c_float),
        ("eps2d", C.c_float), ("w_l2", C.c_float), ("w_l1", C.c_float), ("w_ssim", C.c_float),
        ("lr_pos", C.c_float), ("lr_sh", C.c_float), ("lr_opacity", C.c_float),
        ("lr_scale", C.c_float), ("lr_rot", C.c_float),
        ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
        ("sh_update_interval", C.c_uint32), ("pos_lr_schedule", C.c_uint32),
        ("render_f16", C.c_uint32), ("importance_decay", C.c_float),
        ("persistent", C.c_uint32), ("use_graph", C.c_uint32), ("keep_grads", C.c_uint32),
        ("plain_wf_flags", C.c_uint32), ("xcd_remap", C.c_uint32), ("binned_build", C.c_uint32),
        ("geom_flags", C.c_uint32), ("split_backward", C.c_uint32), ("tile_order", C.c_uint32),
    ]


class DensifyRule(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("grad_threshold", "grad_abs_threshold", "size_threshold", "min_visibility",
                                          "importance_threshold", "pixel_scale")]


def densify_rule(scene_extent: float, train_width: int, **kw) -> DensifyRule:
    """The reference trainer's densify thresholds (DEFAULT_DENSIFICATION_CONFIG,
    densification.ts:26-33; size threshold = scene extent x 0.001, pixel
    scale = training width / 2; importance threshold 5), overridable by
    keyword."""
    r = DensifyRule(grad_threshold=0.003, grad_abs_threshold=0.015, size_threshold=scene_extent * 0.001,
                    min_visibility=25.0, importance_threshold=5.0, pixel_scale=max(train_width, 1) / 2.0)
    for k, v in kw.items():
        setattr(r, k, v)
    return r


class StepStats(C.Structure):
    _fields_ = [("n_isect", C.c_uint64), ("n_views", C.c_uint32), ("sort_passes", C.c_uint32),
                ("key_bits", C.c_uint32), ("binned", C.c_uint32), ("l1_sum", C.c_double)]


class TuneEntry(C.Structure):
    _fields_ = [("batch", C.c_uint32), ("persistent", C.c_uint32), ("xcd_remap", C.c_uint32),
                ("tile_size", C.c_uint32), ("split_backward", C.c_uint32), ("ms_per_step", C.c_double)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("total_ms", C.c_double), ("launches", C.c_uint64)]


_P = C.POINTER(C.c_float)
_U32 = C.POINTER(C.c_uint32)
_I32 = C.POINTER(C.c_int32)
_U8 = C.POINTER(C.c_uint8)
_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "sv_version": (C.c_char_p, []),
    "sv_default_config": (C.c_int, [C.POINTER(Config)]),
    "sv_create": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]),
    "sv_destroy": (C.c_int, [C.c_void_p]),
    "sv_last_error": (C.c_char_p, [C.c_void_p]),
    "sv_set_config": (C.c_int, [C.c_void_p, C.POINTER(Config)]),
    "sv_get_config": (C.c_int, [C.c_void_p, C.POINTER(Config)]),
    "sv_set_gaussians": (C.c_int, [C.c_void_p, C.c_size_t, _P, _P, _P, _P, _P]),
    "sv_get_gaussians": (C.c_int, [C.c_void_p, _P, _P, _P, _P, _P]),
    "sv_get_count": (C.c_size_t, [C.c_void_p]),
    "sv_get_moments": (C.c_int, [C.c_void_p, C.c_int, _P, _P, _P, _P, _P]),
    "sv_set_moments": (C.c_int, [C.c_void_p, C.c_int, _P, _P, _P, _P, _P]),
    "sv_get_step": (C.c_int, [C.c_void_p, _U32]),
    "sv_set_step": (C.c_int, [C.c_void_p, C.c_uint32]),
    "sv_set_views": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Camera)]),
    "sv_set_target": (C.c_int, [C.c_void_p, C.c_int, _P]),
    "sv_build_tiles": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint64)]),
    "sv_forward": (C.c_int, [C.c_void_p, C.c_int]),
    "sv_loss": (C.c_int, [C.c_void_p, C.c_int]),
    "sv_set_residual": (C.c_int, [C.c_void_p, C.c_int, _P]),
    "sv_backward": (C.c_int, [C.c_void_p, C.c_int]),
    "sv_backward_fixed": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32]),
    "sv_zero_grad": (C.c_int, [C.c_void_p]),
    "sv_adam": (C.c_int, [C.c_void_p]),
    "sv_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(StepStats)]),
    "sv_render_batched": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int]),
    "sv_render_camera": (C.c_int, [C.c_void_p, C.POINTER(Camera), _P, _P]),
    "sv_render_camera_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Camera), _P, _P]),
    "sv_read_image": (C.c_int, [C.c_void_p, C.c_int, _P]),
    "sv_read_alpha": (C.c_int, [C.c_void_p, C.c_int, _P]),
    "sv_read_last_ids": (C.c_int, [C.c_void_p, C.c_int, _U32]),
    "sv_read_residual": (C.c_int, [C.c_void_p, C.c_int, _P]),
    "sv_read_keys": (C.c_int, [C.c_void_p, C.c_int, _U32, _U32]),
    "sv_read_tile_offsets": (C.c_int, [C.c_void_p, C.c_int, _U32]),
    "sv_read_projections": (C.c_int, [C.c_void_p, C.c_int, _P]),
    "sv_read_grads": (C.c_int, [C.c_void_p, _P, _P, _P, _P, _P, _P]),
    "sv_read_tile_grads_fixed": (C.c_int, [C.c_void_p, C.c_int, _I32]),
    "sv_read_stats": (C.c_int, [C.c_void_p, _P, _P]),
    "sv_prune_mask": (C.c_int, [C.c_void_p, _U8, C.POINTER(C.c_size_t)]),
    "sv_prune_opacity": (C.c_int, [C.c_void_p, C.c_float, C.POINTER(C.c_size_t)]),
    "sv_densify": (C.c_int, [C.c_void_p, _U8, C.c_uint64, C.POINTER(C.c_size_t)]),
    "sv_reset_opacity": (C.c_int, [C.c_void_p, C.c_float, C.POINTER(C.c_size_t)]),
    "sv_knn_init": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p]),
    "sv_multiview_scores": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int, _U32, _P, C.POINTER(C.c_int)]),
    "sv_densify_decide": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_float, C.c_float, C.c_size_t,
                                    _U8, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "sv_densify_decide_grad": (C.c_int, [C.c_void_p, C.POINTER(DensifyRule), _U32, C.c_uint32, C.c_size_t, _U8,
                                         C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "sv_read_tile_l1": (C.c_int, [C.c_void_p, C.c_int, _P]),
    "sv_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(TuneEntry),
                              C.POINTER(C.c_int)]),
    "sv_save_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p]),
    "sv_load_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p]),
    "sv_memory_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "sv_reset_memory_peak": (C.c_int, [C.c_void_p]),
    "sv_enable_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sv_timings": (C.c_int, [C.c_void_p, C.POINTER(KernelTime), C.POINTER(C.c_int)]),
    "sv_reset_timings": (C.c_int, [C.c_void_p]),
    "sv_sync": (C.c_int, [C.c_void_p]),
    "sv_scene_gaussians": (C.c_int, [C.c_uint64, C.c_size_t, C.c_float, _P, _P, _P, _P, _P]),
    "sv_scene_jitter": (C.c_int, [C.c_uint64, C.c_size_t, C.c_float, _P]),
    "sv_scene_camera": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, C.c_float, C.POINTER(Camera)]),
    "sv_scene_uniform": (C.c_double, [C.c_uint64, C.c_uint64]),
    "sv_ingest_error": (C.c_char_p, []),
    "sv_ply_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_size_t), C.POINTER(C.c_uint32)]),
    "sv_ply_load": (C.c_int, [C.c_char_p, C.c_size_t, _P, _P, _P, _P, _P, _P]),
    "sv_ply_save": (C.c_int, [C.c_char_p, C.c_size_t, _P, _P, _P, _P, _P, C.c_uint32]),
    "sv_colmap_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "sv_colmap_load": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(Camera), C.c_char_p, C.c_size_t, C.c_size_t,
                                 _P, _U8]),
    "sv_points_to_gaussians": (C.c_int, [C.c_size_t, _P, _U8, _P, _P, _P, _P, _P, _P]),
    "sv_scene_normal": (C.c_double, [C.c_uint64, C.c_uint64]),
}

SYMBOLS = tuple(_SIGS)


def lib():
    """Load libsv.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run python skyfall-viewer_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class SvError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"sv status {status}: {msg}")
        self.status = status


def _fp(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def default_config() -> Config:
    c = Config()
    lib().sv_default_config(C.byref(c))
    return c


# ----------------------------------------------------------------------
# seeded synthetic scenes (host side of libsv)
# ----------------------------------------------------------------------
def scene_gaussians(seed: int, n: int, half_extent: float = 1.0) -> dict:
    pos = np.empty((n, 3), np.float32)
    ls = np.empty((n, 3), np.float32)
    q = np.empty((n, 4), np.float32)
    lg = np.empty((n,), np.float32)
    sh = np.empty((n, 48), np.float32)
    st = lib().sv_scene_gaussians(seed, n, half_extent, _fp(pos), _fp(ls), _fp(q), _fp(lg), _fp(sh))
    if st:
        raise SvError(st, "sv_scene_gaussians")
    return {"pos": pos, "logscale": ls, "quat": q, "logit": lg, "sh": sh}


def _ing(st, what):
    if st:
        raise SvError(st, f"{what}: {lib().sv_ingest_error().decode()}")


def _param_arrays(n):
    return {"pos": np.zeros((n, 3), np.float32), "logscale": np.zeros((n, 3), np.float32),
            "quat": np.zeros((n, 4), np.float32), "logit": np.zeros((n,), np.float32),
            "sh": np.zeros((n, 48), np.float32)}


def ply_load(path: str):
    """(Gaussians dict, SH degree, scene extent) of a PLY file."""
    n, deg = C.c_size_t(), C.c_uint32()
    _ing(lib().sv_ply_info(os.fsencode(path), C.byref(n), C.byref(deg)), "ply_info")
    g = _param_arrays(n.value)
    ext = np.zeros(1, np.float32)
    _ing(lib().sv_ply_load(os.fsencode(path), n.value, *[_fp(g[k]) for k in ("pos", "logscale", "quat", "logit", "sh")],
                           _fp(ext)), "ply_load")
    return g, deg.value, float(ext[0])


def ply_save(path: str, g: dict, sh_deg: int = 3):
    arrs = [_f32(g[k]) for k in ("pos", "logscale", "quat", "logit", "sh")]
    _ing(lib().sv_ply_save(os.fsencode(path), arrs[0].shape[0], *[_fp(a) for a in arrs], sh_deg), "ply_save")


def colmap_load(sparse_dir: str):
    """(cameras, image names, xyz float32[n, 3], rgb uint8[n, 3]) of a COLMAP sparse model."""
    ni, npt = C.c_size_t(), C.c_size_t()
    _ing(lib().sv_colmap_info(os.fsencode(sparse_dir), C.byref(ni), C.byref(npt)), "colmap_info")
    cams = (Camera * max(ni.value, 1))()
    names = C.create_string_buffer(max(ni.value, 1) * 256)
    xyz = np.zeros((npt.value, 3), np.float32)
    rgb = np.zeros((npt.value, 3), np.uint8)
    _ing(lib().sv_colmap_load(os.fsencode(sparse_dir), ni.value, cams, names, 256, npt.value, _fp(xyz),
                              rgb.ctypes.data_as(_U8)), "colmap_load")
    raw = names.raw
    nm = [raw[i * 256:(i + 1) * 256].split(b"\0", 1)[0].decode() for i in range(ni.value)]
    return list(cams)[:ni.value], nm, xyz, rgb


def points_to_gaussians(xyz: np.ndarray, rgb: np.ndarray):
    x = _f32(xyz)
    c = np.ascontiguousarray(rgb, np.uint8)
    g = _param_arrays(x.shape[0])
    ext = np.zeros(1, np.float32)
    _ing(lib().sv_points_to_gaussians(x.shape[0], _fp(x), c.ctypes.data_as(_U8),
                                      *[_fp(g[k]) for k in ("pos", "logscale", "quat", "logit", "sh")], _fp(ext)),
         "points_to_gaussians")
    return g, float(ext[0])


def scene_jitter(seed: int, pos: np.ndarray, sigma: float) -> np.ndarray:
    out = np.array(pos, dtype=np.float32, copy=True, order="C")
    st = lib().sv_scene_jitter(seed, out.shape[0], sigma, _fp(out))
    if st:
        raise SvError(st, "sv_scene_jitter")
    return out


def scene_camera(eye, target, w: int, h: int, fovy_deg: float = 60.0) -> Camera:
    cam = Camera()
    e = np.asarray(eye, np.float32)
    t = np.asarray(target, np.float32)
    st = lib().sv_scene_camera(_fp(e), _fp(t), w, h, fovy_deg, C.byref(cam))
    if st:
        raise SvError(st, "sv_scene_camera")
    return cam


def camera_arrays(cam: Camera) -> dict:
    return {k: np.array(getattr(cam, k), np.float32) for k in ("view", "view_inv", "proj", "proj_inv",
                                                             "viewport", "focal")}


def ring_cameras(n: int, radius: float, w: int, h: int, fovy: float = 60.0):
    """Cameras on a ring in the y=0 plane at azimuth k*360/n, looking at the origin."""
    cams = []
    for k in range(n):
        az = 2.0 * np.pi * k / n
        cams.append(scene_camera([radius * np.sin(az), 0.0, radius * np.cos(az)], [0, 0, 0], w, h, fovy))
    return cams


def sphere_cameras(seed: int, n: int, radius: float, wmin=320, wmax=1920, hmin=240, hmax=1080,
                   fovy: float = 60.0):
    """Cameras uniform on a sphere with seeded uneven resolutions."""
    L = lib()
    cams, sizes = [], []
    for k in range(n):
        z = 2.0 * L.sv_scene_uniform(seed, 8 * k) - 1.0
        phi = 2.0 * np.pi * L.sv_scene_uniform(seed, 8 * k + 1)
        r = np.sqrt(max(0.0, 1.0 - z * z))
        eye = [radius * r * np.cos(phi), radius * z, radius * r * np.sin(phi)]
        w = wmin + int(L.sv_scene_uniform(seed, 8 * k + 2) * (wmax - wmin + 1))
        h = hmin + int(L.sv_scene_uniform(seed, 8 * k + 3) * (hmax - hmin + 1))
        cams.append(scene_camera(eye, [0, 0, 0], w, h, fovy))
        sizes.append((w, h))
    return cams, sizes


# ----------------------------------------------------------------------
# context
# ----------------------------------------------------------------------
class Context:
    """One libsv context (one HIP stream on one device)."""

    def __init__(self, capacity: int, device: int = 0):
        h = C.c_void_p()
        st = lib().sv_create(device, capacity, C.byref(h))
        if st:
            raise SvError(st, lib().sv_last_error(None).decode())
        self.h = h
        self.views: list[tuple[int, int]] = []

    def close(self):
        if self.h:
            lib().sv_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, st, what=""):
        if st:
            raise SvError(st, f"{what}: {lib().sv_last_error(self.h).decode()}")

    # configuration ----------------------------------------------------
    def config(self) -> Config:
        c = Config()
        self._ck(lib().sv_get_config(self.h, C.byref(c)), "get_config")
        return c

    def set_config(self, cfg: Config = None, **kw):
        c = cfg if cfg is not None else self.config()
        names = {f[0] for f in Config._fields_}
        for k, v in kw.items():
            if k not in names:
                raise AttributeError(f"sv_config has no field {k!r}")
            if k == "bg":
                for i in range(3):
                    c.bg[i] = v[i]
            else:
                setattr(c, k, v)
        self._ck(lib().sv_set_config(self.h, C.byref(c)), "set_config")

    # parameters ---------------------------------------------------------
    def set_gaussians(self, g: dict):
        n = g["pos"].shape[0]
        arrs = [_f32(g[k]) for k in ("pos", "logscale", "quat", "logit", "sh")]
        self._ck(lib().sv_set_gaussians(self.h, n, *[_fp(a) for a in arrs]), "set_gaussians")

    @property
    def count(self) -> int:
        return lib().sv_get_count(self.h)

    def _alloc_params(self):
        n = self.count
        return {"pos": np.empty((n, 3), np.float32), "logscale": np.empty((n, 3), np.float32),
                "quat": np.empty((n, 4), np.float32), "logit": np.empty((n,), np.float32),
                "sh": np.empty((n, 48), np.float32)}

    def get_gaussians(self) -> dict:
        o = self._alloc_params()
        self._ck(lib().sv_get_gaussians(self.h, *[_fp(o[k]) for k in ("pos", "logscale", "quat", "logit", "sh")]))
        return o

    def get_moments(self, which: int) -> dict:
        o = self._alloc_params()
        self._ck(lib().sv_get_moments(self.h, which, *[_fp(o[k]) for k in ("pos", "logscale", "quat", "logit", "sh")]))
        return o

    def set_moments(self, which: int, m: dict):
        arrs = [_f32(m[k]) for k in ("pos", "logscale", "quat", "logit", "sh")]
        self._ck(lib().sv_set_moments(self.h, which, *[_fp(a) for a in arrs]), "set_moments")

    @property
    def step(self) -> int:
        t = C.c_uint32()
        self._ck(lib().sv_get_step(self.h, C.byref(t)))
        return t.value

    # views --------------------------------------------------------------
    def set_views(self, cams):
        arr = (Camera * len(cams))(*cams)
        self._ck(lib().sv_set_views(self.h, len(cams), arr), "set_views")
        self.views = [(int(round(c.viewport[0])), int(round(c.viewport[1]))) for c in cams]

    def set_target(self, v: int, rgb: np.ndarray):
        w, h = self.views[v]
        a = _f32(rgb, (h, w, 3))
        self._ck(lib().sv_set_target(self.h, v, _fp(a)), "set_target")

    # function-style pieces ------------------------------------------------
    def build_tiles(self, v: int) -> int:
        n = C.c_uint64()
        self._ck(lib().sv_build_tiles(self.h, v, C.byref(n)), "build_tiles")
        return n.value

    def forward(self, v: int):
        self._ck(lib().sv_forward(self.h, v), "forward")

    def loss(self, v: int):
        self._ck(lib().sv_loss(self.h, v), "loss")

    def set_residual(self, v: int, rgb: np.ndarray):
        w, h = self.views[v]
        a = _f32(rgb, (h, w, 3))
        self._ck(lib().sv_set_residual(self.h, v, _fp(a)), "set_residual")

    def backward(self, v: int):
        self._ck(lib().sv_backward(self.h, v), "backward")

    def backward_fixed(self, v: int, flags: int = 0):
        self._ck(lib().sv_backward_fixed(self.h, v, flags), "backward_fixed")

    def zero_grad(self):
        self._ck(lib().sv_zero_grad(self.h), "zero_grad")

    def adam(self):
        self._ck(lib().sv_adam(self.h), "adam")

    def step_views(self, views, mode: int = MODE_PLAIN, stats: bool = False):
        vs = (C.c_int * len(views))(*views)
        s = StepStats() if stats else None
        self._ck(lib().sv_step(self.h, vs, len(views), mode, C.byref(s) if s is not None else None), "step")
        return s

    def render_camera(self, cam: Camera):
        """Render any camera through the tile pipeline: (rgb H x W x 3, alpha H x W)."""
        w, h = int(round(cam.viewport[0])), int(round(cam.viewport[1]))
        rgb = np.empty((h, w, 3), np.float32)
        alpha = np.empty((h, w), np.float32)
        self._ck(lib().sv_render_camera(self.h, C.byref(cam), _fp(rgb), _fp(alpha)), "render_camera")
        return rgb, alpha

    def render_camera_noalloc(self, cam: Camera):
        """render_camera through the fixed device scratch (no allocation)."""
        w, h = int(round(cam.viewport[0])), int(round(cam.viewport[1]))
        rgb = np.empty((h, w, 3), np.float32)
        alpha = np.empty((h, w), np.float32)
        self._ck(lib().sv_render_camera_noalloc(self.h, C.byref(cam), _fp(rgb), _fp(alpha)), "render_camera_noalloc")
        return rgb, alpha

    def render_batched(self, views):
        vs = (C.c_int * len(views))(*views)
        self._ck(lib().sv_render_batched(self.h, vs, len(views)), "render_batched")

    # readbacks --------------------------------------------------------------
    def image(self, v: int) -> np.ndarray:
        w, h = self.views[v]
        o = np.empty((h, w, 3), np.float32)
        self._ck(lib().sv_read_image(self.h, v, _fp(o)), "read_image")
        return o

    def residual(self, v: int) -> np.ndarray:
        w, h = self.views[v]
        o = np.empty((h, w, 3), np.float32)
        self._ck(lib().sv_read_residual(self.h, v, _fp(o)), "read_residual")
        return o

    def alpha(self, v: int) -> np.ndarray:
        w, h = self.views[v]
        o = np.empty((h, w), np.float32)
        self._ck(lib().sv_read_alpha(self.h, v, _fp(o)), "read_alpha")
        return o

    def last_ids(self, v: int) -> np.ndarray:
        w, h = self.views[v]
        o = np.empty((h, w), np.uint32)
        self._ck(lib().sv_read_last_ids(self.h, v, o.ctypes.data_as(_U32)), "read_last_ids")
        return o

    def keys(self, v: int, n: int):
        k = np.empty(n, np.uint32)
        val = np.empty(n, np.uint32)
        self._ck(lib().sv_read_keys(self.h, v, k.ctypes.data_as(_U32), val.ctypes.data_as(_U32)), "read_keys")
        return k, val

    def tile_offsets(self, v: int) -> np.ndarray:
        w, h = self.views[v]
        ts = self.config().tile_size
        nt = ((w + ts - 1) // ts) * ((h + ts - 1) // ts)
        o = np.empty(nt + 1, np.uint32)
        self._ck(lib().sv_read_tile_offsets(self.h, v, o.ctypes.data_as(_U32)), "read_tile_offsets")
        return o

    def projections(self, v: int) -> np.ndarray:
        o = np.empty((self.count, 12), np.float32)
        self._ck(lib().sv_read_projections(self.h, v, _fp(o)), "read_projections")
        return o

    def grads(self) -> dict:
        o = self._alloc_params()
        o["geom"] = np.empty((self.count, 5), np.float32)
        self._ck(lib().sv_read_grads(self.h, *[_fp(o[k]) for k in ("pos", "logscale", "quat", "logit", "sh", "geom")]))
        return o

    def tile_grads_fixed(self, v: int, n: int) -> np.ndarray:
        o = np.empty((n, 54), np.int32)
        self._ck(lib().sv_read_tile_grads_fixed(self.h, v, o.ctypes.data_as(_I32)), "read_tile_grads_fixed")
        return o

    def stats(self):
        imp = np.empty((self.count, 4), np.float32)
        acc = np.empty((self.count, 4), np.float32)
        self._ck(lib().sv_read_stats(self.h, _fp(imp), _fp(acc)))
        return imp, acc

    # prune / densify --------------------------------------------------------
    def prune_mask(self, remove: np.ndarray) -> int:
        r = np.ascontiguousarray(remove, dtype=np.uint8)
        n = C.c_size_t()
        self._ck(lib().sv_prune_mask(self.h, r.ctypes.data_as(_U8), C.byref(n)), "prune_mask")
        return n.value

    def prune_opacity(self, min_opacity: float) -> int:
        n = C.c_size_t()
        self._ck(lib().sv_prune_opacity(self.h, min_opacity, C.byref(n)), "prune_opacity")
        return n.value

    def densify(self, actions: np.ndarray, seed: int) -> int:
        a = np.ascontiguousarray(actions, dtype=np.uint8)
        n = C.c_size_t()
        self._ck(lib().sv_densify(self.h, a.ctypes.data_as(_U8), seed, C.byref(n)), "densify")
        return n.value

    def save_checkpoint(self, path: str):
        self._ck(lib().sv_save_checkpoint(self.h, os.fsencode(path)), "save_checkpoint")

    def load_checkpoint(self, path: str):
        self._ck(lib().sv_load_checkpoint(self.h, os.fsencode(path)), "load_checkpoint")

    def autotune(self, views, steps: int = 3) -> list:
        """Time every tuner candidate on `views`; the fastest stays configured."""
        vs = (C.c_int * len(views))(*views)
        n = C.c_int(0)
        self._ck(lib().sv_autotune(self.h, vs, len(views), steps, None, C.byref(n)), "autotune")
        arr = (TuneEntry * n.value)()
        self._ck(lib().sv_autotune(self.h, vs, len(views), steps, arr, C.byref(n)), "autotune")
        return [{"batch": e.batch, "persistent": e.persistent, "xcd_remap": e.xcd_remap, "tile": e.tile_size,
                 "split_backward": e.split_backward, "ms_per_step": e.ms_per_step} for e in arr]

    def reset_opacity(self, max_opacity: float = 0.1) -> int:
        n = C.c_size_t()
        self._ck(lib().sv_reset_opacity(self.h, max_opacity, C.byref(n)), "reset_opacity")
        return n.value

    def knn_init(self, scene_extent: float) -> np.ndarray:
        """Density-aware initial log-scales (initializeKnn); returns the
        per-Gaussian mean distance to the three nearest neighbours."""
        dens = np.zeros(self.count, np.float32)
        self._ck(lib().sv_knn_init(self.h, scene_extent, dens.ctypes.data if dens.size else None), "knn_init")
        return dens

    def multiview_scores(self, views):
        """generateMultiViewScores over `views` in order: (importance u32[N],
        pruning f32[N], views scored)."""
        vs = (C.c_int * len(views))(*views)
        imp = np.zeros(self.count, np.uint32)
        pru = np.zeros(self.count, np.float32)
        ns = C.c_int()
        self._ck(lib().sv_multiview_scores(self.h, vs, len(views), imp.ctypes.data_as(_U32), _fp(pru),
                                           C.byref(ns)), "multiview_scores")
        return imp, pru, ns.value

    def densify_decide(self, views, percentile: float = 0.9, split_px: float = 8.0, max_new: int = 0):
        """Returns (actions uint8[N], n_clone, n_split); max_new > 0 keeps only
        the max_new candidates of highest score (the reference's cap is
        max(1, N // 10))."""
        vs = (C.c_int * len(views))(*views)
        act = np.zeros(self.count, np.uint8)
        nc, ns = C.c_size_t(), C.c_size_t()
        self._ck(lib().sv_densify_decide(self.h, vs, len(views), percentile, split_px, max_new,
                                         act.ctypes.data_as(_U8), C.byref(nc), C.byref(ns)), "densify_decide")
        return act, nc.value, ns.value

    def densify_decide_grad(self, rule: "DensifyRule", importance: np.ndarray, n_samples: int, max_new: int = 0):
        """The reference trainer's densify decision from the step statistics and
        multi-view importance counts. Returns (actions uint8[N], n_clone,
        n_split)."""
        imp = np.ascontiguousarray(importance, np.uint32)
        if imp.shape != (self.count,):
            raise ValueError(f"importance must hold {self.count} counts")
        act = np.zeros(self.count, np.uint8)
        nc, ns = C.c_size_t(), C.c_size_t()
        self._ck(lib().sv_densify_decide_grad(self.h, C.byref(rule), imp.ctypes.data_as(_U32), n_samples, max_new,
                                              act.ctypes.data_as(_U8), C.byref(nc), C.byref(ns)),
                 "densify_decide_grad")
        return act, nc.value, ns.value

    def tile_l1(self, v: int) -> np.ndarray:
        w, h = self.views[v]
        out = np.zeros(((w + 15) // 16) * ((h + 15) // 16), np.float32)
        self._ck(lib().sv_read_tile_l1(self.h, v, _fp(out)), "read_tile_l1")
        return out

    # device memory ------------------------------------------------------------
    def memory_stats(self):
        """(bytes of device memory held now, peak since creation or the last reset)."""
        a, b = C.c_uint64(), C.c_uint64()
        self._ck(lib().sv_memory_stats(self.h, C.byref(a), C.byref(b)), "memory_stats")
        return a.value, b.value

    def reset_memory_peak(self):
        self._ck(lib().sv_reset_memory_peak(self.h), "reset_memory_peak")

    # timing -------------------------------------------------------------------
    def enable_timing(self, on: bool = True):
        self._ck(lib().sv_enable_timing(self.h, 1 if on else 0))

    def reset_timings(self):
        self._ck(lib().sv_reset_timings(self.h))

    def timings(self) -> dict:
        n = C.c_int(64)
        arr = (KernelTime * 64)()
        self._ck(lib().sv_timings(self.h, arr, C.byref(n)))
        return {arr[i].name.decode(): (arr[i].total_ms, arr[i].launches) for i in range(min(n.value, 64))}

    def sync(self):
        self._ck(lib().sv_sync(self.h))
