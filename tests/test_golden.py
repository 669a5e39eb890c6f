"""Golden fixtures (tests/golden, written by tests/golden/make_golden.py):
the numpy restatement of the seeded generator equals the library's C++
generator bit for bit, and the oracle still reproduces every recorded
digest (CPU only)."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as mg  # noqa: E402

with open(os.path.join(HERE, "golden", "digests.json")) as fh:
    MANIFEST = json.load(fh)


@pytest.mark.parametrize("seed,n,half", [(0, 4096, 1.0), (5, 1000, 2.0), (123456789, 77, 1.0)])
def test_numpy_generator_equals_library_generator(sv, seed, n, half):
    a = mg.gen_gaussians(seed, n, half)
    b = sv.scene_gaussians(seed, n, half)
    for k in a:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    ja = mg.gen_jitter(seed + 1, a["pos"], 0.02)
    jb = sv.scene_jitter(seed + 1, b["pos"], 0.02)
    assert np.array_equal(ja.view(np.uint32), jb.view(np.uint32))


@pytest.mark.parametrize("eye,w,h", [((0, 0, 3), 256, 256), ((2.1, -0.4, 1.7), 1920, 1080), ((0, 3, 0), 64, 48)])
def test_numpy_camera_equals_library_camera(sv, eye, w, h):
    a = mg.gen_camera(eye, (0, 0, 0), w, h, 60.0)
    b = sv.camera_arrays(sv.scene_camera(list(eye), [0, 0, 0], w, h, 60.0))
    for k in ("view", "proj", "viewport", "focal"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_generator_inputs_match_digests(name):
    n = mg.SCENES[name][0]
    g = mg.gen_gaussians(0, n, 1.0)
    for k, d in MANIFEST[name]["inputs"].items():
        assert mg.digest(g[k]) == d, k


@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_oracle_reproduces_golden(orc, name):
    g, target, r = mg.run_oracle(name)
    ent = MANIFEST[name]
    assert r["n_isect"] == ent["n_isect"]
    assert mg.digest(target) == ent["target"]
    for k in mg.ARRAYS:
        assert mg.digest(r[k]) == ent["outputs"][k], k
    for k in mg.GRADS:
        assert mg.digest(r["grads"][k]) == ent["grads"][k], k
    if name == "adjoint64":
        z = np.load(os.path.join(HERE, "golden", "adjoint64_4096_64x64.npz"))
        for k in mg.ARRAYS:
            assert np.array_equal(z[k], r[k]), k
