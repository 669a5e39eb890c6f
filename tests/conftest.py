"""Shared test setup: import paths, the `gpu` marker and seeded scenes."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("skyfall-viewer_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsv.so on the device)")


@pytest.fixture(scope="session")
def sv():
    import sv as _sv
    _sv.lib()
    return _sv


@pytest.fixture(scope="session")
def orc():
    import oracle as _orc
    _orc.lib()
    return _orc


def small_scene(sv, n=16384, w=256, h=256, seed=0, jitter_seed=1):
    """The seeded small scene: n Gaussians in [-1,1]^3, one camera at
    (0,0,3) looking at the origin with a 60 degree vertical FOV."""
    g = sv.scene_gaussians(seed, n, 1.0)
    cam = sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], w, h, 60.0)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(jitter_seed, g["pos"], 0.02)
    return g, tg, cam


def close_rel(a, b, rtol):
    """max |a-b| / max(|b|) (scale-normalised error of a tensor)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return float(np.abs(a - b).max() / scale)
