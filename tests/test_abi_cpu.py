"""The C ABI library loads without a GPU, exports every symbol the header
declares, and the host-only harness functions behave (no device calls)."""
import ctypes
import os
import re

import numpy as np

from conftest import ROOT


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "sv_api.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sv_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols(sv):
    L = sv.lib()
    syms = _header_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding declares exactly the header's functions
    assert sorted(sv.SYMBOLS) == syms


def test_struct_sizes(sv):
    assert ctypes.sizeof(sv.Camera) == 272  # CameraUniforms, common.wgsl:21-28
    cfg = sv.default_config()
    assert cfg.tile_size == 16 and cfg.batch == 64
    assert abs(cfg.eps2d - 0.3) < 1e-7 and cfg.radius_clip == 512.0
    assert (cfg.w_l2, round(cfg.w_l1, 6), round(cfg.w_ssim, 6)) == (0.0, 0.8, 0.2)


def test_create_without_gpu_fails_cleanly(sv):
    import torch

    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    st = sv.lib().sv_create(0, 1024, ctypes.byref(h))
    assert st in (sv.SV_E_HIP, sv.SV_E_ARG)
    assert not h.value
    assert sv.lib().sv_last_error(None)


def test_scene_generator_is_deterministic_and_f16(sv, orc):
    a = sv.scene_gaussians(0, 1000, 1.0)
    b = sv.scene_gaussians(0, 1000, 1.0)
    c = sv.scene_gaussians(1, 1000, 1.0)
    for k in a:
        assert np.array_equal(a[k], b[k])
        assert not np.array_equal(a[k], c[k])
        # every value is f16-representable (rounded to nearest even)
        assert np.array_equal(a[k].astype(np.float16).astype(np.float32), a[k])
    # distributions (loose moments)
    big = sv.scene_gaussians(3, 20000, 1.0)
    assert abs(big["pos"].mean()) < 0.02 and abs(big["pos"].max() - 1.0) < 1e-3
    assert abs(big["logscale"].mean() + 4.5) < 0.02 and abs(big["logscale"].std() - 0.3) < 0.02
    assert abs(big["logit"].std() - 1.0) < 0.03
    assert abs(big["sh"][:, :3].mean() - 0.5) < 0.01 and abs(big["sh"][:, 3:].std() - 0.05) < 0.005
    assert np.allclose(np.linalg.norm(big["quat"], axis=1), 1.0, atol=2e-3)


def test_camera_matches_pinhole_formula(sv):
    cam = sv.scene_camera([0, 0, 3], [0, 0, 0], 256, 256, 60.0)
    f = 128.0 / np.tan(np.radians(30.0))
    assert abs(cam.focal[0] - f) < 1e-3 and abs(cam.focal[1] - f) < 1e-3
    P = np.array(cam.proj).reshape(4, 4).T  # column-major -> row-major
    # get_pinhole_projection_matrix (training-camera.ts:29-52) after transpose
    exp = np.array([[2 * f / 256, 0, 0, 0], [0, -2 * f / 256, 0, 0], [0, 0, 1, 0], [0, 0, 1, 0]])
    assert np.allclose(P, exp, atol=1e-6)
    V = np.array(cam.view).reshape(4, 4).T
    # the eye maps to the camera origin, the target to +z
    assert np.allclose(V @ [0, 0, 3, 1], [0, 0, 0, 1], atol=1e-6)
    assert np.allclose(V @ [0, 0, 0, 1], [0, 0, 3, 1], atol=1e-6)
    Vi = np.array(cam.view_inv).reshape(4, 4).T
    assert np.allclose(Vi @ V, np.eye(4), atol=1e-6)
